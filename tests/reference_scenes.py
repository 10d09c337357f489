"""The small closed-form scenes of the reference's estimator tests
(crates/cenote/src/render/tests.rs), built as plain scene data so the
oracle (CPU) and the HIP path (GPU) render the very same inputs."""

import numpy as np

from cenote_amd import _abi
from cenote_amd.scene import CameraData, Material, MeshData, OpenPBR, SceneData, constant_environment, identity
from cenote_amd.shapes import ground_plane


def delta_light_scene(kind: int, vector, color) -> SceneData:
    """render/tests.rs:293-343: a white Lambert plane under a black sky, lit
    by exactly one delta light; camera just above, looking obliquely down."""
    s = SceneData()
    white = OpenPBR(base_color=(1.0, 1.0, 1.0), specular_weight=0.0).lower()
    s.add_instance(s.add_mesh(ground_plane(5.0)), identity(), white)
    s.delta_lights.append((kind, tuple(vector), tuple(color)))
    s.env_texels = constant_environment((0.0, 0.0, 0.0))
    s.camera = CameraData(position=(0.0, 1.0, 0.0), look_at=(0.0, 0.0, -1.0))
    return s


def distant_light_scene() -> SceneData:
    return delta_light_scene(_abi.LIGHT_DISTANT, (0.0, -1.0, 0.0), (np.pi,) * 3)


def point_light_scene() -> SceneData:
    return delta_light_scene(_abi.LIGHT_POINT, (0.0, 1000.0, 0.0), (np.pi * 1e6,) * 3)


def aov_quad_scene() -> SceneData:
    """render/tests.rs:1117-1160: a +-0.6 matte quad facing +z at distance 2."""
    s = SceneData()
    p = np.array([[-0.6, -0.6, 0], [0.6, -0.6, 0], [0.6, 0.6, 0], [-0.6, 0.6, 0]], np.float32)
    n = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    mesh = MeshData(p, n, np.array([[0, 1, 2], [0, 2, 3]], np.uint32))
    matte = OpenPBR(base_color=(0.7, 0.2, 0.1), specular_weight=0.0).lower()
    s.add_instance(s.add_mesh(mesh), identity(), matte)
    s.env_texels = constant_environment((0.0, 0.0, 0.0))
    s.camera = CameraData(position=(0.0, 0.0, 2.0), look_at=(0.0, 0.0, 0.0), vfov_degrees=53.1301)
    return s


def agreement_scene() -> SceneData:
    """An MIS-agreement scene (wavefront.rs:1246-1557 style): a glossy floor
    and a diffuse sphere under an area light and a gradient sky, so both
    strategies reach both kinds of light."""
    from cenote_amd.shapes import uv_sphere
    from cenote_amd.scene import trs

    s = SceneData()
    floor = OpenPBR(base_color=(0.6, 0.6, 0.6), specular_roughness=0.3).lower()
    s.add_instance(s.add_mesh(ground_plane(4.0)), identity(), floor)
    ball = OpenPBR(base_color=(0.8, 0.3, 0.2), specular_weight=0.0, base_diffuse_roughness=0.5).lower()
    s.add_instance(s.add_mesh(uv_sphere(24, 12)), trs((0.0, 0.6, 0.0), scale=(0.6, 0.6, 0.6)), ball)
    key = Material.emitter((6.0, 5.0, 4.0))
    s.add_instance(s.add_mesh(ground_plane(0.5)), trs((1.0, 2.5, 0.5), (180.0, 0.0, 0.0)), key)
    h, w = 16, 32
    sky = np.ones((h, w, 4), np.float32)
    v = (np.arange(h) + 0.5) / h
    sky[..., 0] = (0.2 + 0.6 * (1 - v))[:, None]
    sky[..., 1] = (0.3 + 0.5 * (1 - v))[:, None]
    sky[..., 2] = (0.5 + 0.4 * (1 - v))[:, None]
    s.env_texels = sky
    s.camera = CameraData(position=(0.0, 1.5, 4.0), look_at=(0.0, 0.5, 0.0), vfov_degrees=40.0)
    return s


def _black_veil(opacity: float) -> Material:
    return OpenPBR(base_color=(0.0, 0.0, 0.0), specular_weight=0.0, geometry_opacity=opacity).lower()


def stacked_planes_scene(planes: int, opacity: float, camera_visible: bool = False) -> SceneData:
    """The distant-light plane of render/tests.rs:352-374 (a white Lambert
    floor lit straight from above, irradiance pi: every pixel's radiance is
    exactly 1) under `planes` black fractional-opacity planes stacked between
    floor and light, so every next-event connection crosses each plane once
    (trace_shadow.slang:52-65): a pixel's radiance is (1 - opacity)^planes,
    or 0 once that falls below 1e-4. The veils are invisible to camera rays
    (camera_visible=False, scene/mod.rs:54-57) unless asked otherwise;
    bounce rays that stop on a veil find a black surface and the sky is
    black, so the next-event term is the whole pixel."""
    s = SceneData()
    white = OpenPBR(base_color=(1.0, 1.0, 1.0), specular_weight=0.0).lower()
    s.add_instance(s.add_mesh(ground_plane(5.0)), identity(), white)
    veil = s.add_mesh(ground_plane(50.0))
    for k in range(planes):
        t = identity().copy()
        t[1, 3] = 0.5 + 0.05 * k
        s.add_instance(veil, t, _black_veil(opacity), camera_visible=camera_visible)
    s.delta_lights.append((_abi.LIGHT_DISTANT, (0.0, -1.0, 0.0), (np.pi,) * 3))
    s.env_texels = constant_environment((0.0, 0.0, 0.0))
    if camera_visible:  # look straight down through the stack
        s.camera = CameraData(position=(0.0, 3.0, 0.0), look_at=(0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0),
                              vfov_degrees=60.0)
    else:
        s.camera = CameraData(position=(0.0, 1.0, 0.0), look_at=(0.0, 0.0, -1.0))
    return s


def invisible_emitter_scene(visible: bool) -> SceneData:
    """scene/prep.rs:467-558: a gray floor (the default material, scaled 8x)
    and an upright emissive panel dead ahead of the camera, with a black sky
    behind; the panel is camera-visible or not."""
    from cenote_amd.scene import trs

    s = SceneData()
    quad = s.add_mesh(ground_plane(1.0))
    s.add_instance(quad, trs(scale=(8.0, 1.0, 8.0)), OpenPBR().lower())
    lamp = OpenPBR(base_color=(0.0, 0.0, 0.0), specular_weight=0.0, emission_luminance=10.0).lower()
    s.add_instance(quad, trs((0.0, 1.0, -1.0), (90.0, 0.0, 0.0), (0.5, 1.0, 0.5)), lamp, camera_visible=visible)
    s.env_texels = constant_environment((0.0, 0.0, 0.0))
    s.camera = CameraData(position=(0.0, 1.0, 4.0), look_at=(0.0, 1.0, 0.0))
    return s


def translucent_scene() -> SceneData:
    """Fractional opacity and the camera mask together in a lit scene: the
    MIS-agreement scene plus a half-opaque glossy sphere in front of the
    ball, a 0.3-opacity colored pane across the light's path and a
    camera-invisible second emitter."""
    from cenote_amd.shapes import uv_sphere
    from cenote_amd.scene import trs

    s = agreement_scene()
    ghost = OpenPBR(base_color=(0.2, 0.6, 0.9), specular_roughness=0.2, geometry_opacity=0.5).lower()
    s.add_instance(s.add_mesh(uv_sphere(24, 12)), trs((-0.7, 0.4, 1.0), scale=(0.4, 0.4, 0.4)), ghost)
    pane = OpenPBR(base_color=(0.9, 0.4, 0.1), specular_weight=0.0, geometry_opacity=0.3).lower()
    s.add_instance(s.add_mesh(ground_plane(0.8)), trs((0.6, 1.6, 0.4)), pane)
    hidden = Material.emitter((2.0, 3.0, 6.0))
    s.add_instance(s.add_mesh(ground_plane(0.4)), trs((-1.2, 2.0, -0.5), (180.0, 0.0, 0.0)), hidden,
                   camera_visible=False)
    return s


WHITE = (1.0, 1.0, 1.0)

# render/tests.rs:427-469: the white-furnace matrix, configuration for
# configuration, built with the reference's Material builders
# (material.rs:118-245).
FURNACE_CONFIGS = {
    "metal r=0.05": Material.metal(WHITE, 0.05),
    "metal r=0.5": Material.metal(WHITE, 0.5),
    "metal r=1.0": Material.metal(WHITE, 1.0),
    "glossy-diffuse r=0.05": Material.glossy(WHITE, 0.0, 0.05),
    "glossy-diffuse r=0.5": Material.glossy(WHITE, 0.0, 0.5),
    "glossy-diffuse r=1.0, rough base": Material.glossy(WHITE, 1.0, 1.0),
    "half metal": Material.glossy(WHITE, 0.0, 0.5).with_metalness(0.5),
    "glossy ior=2.5": Material.glossy(WHITE, 0.0, 0.5).with_ior(2.5),
    "glossy ior=1.1": Material.glossy(WHITE, 0.0, 0.8).with_ior(1.1),
    "coat over diffuse": Material.glossy(WHITE, 0.0, 0.5).with_coat(1.0, 0.3),
    "coat over metal": Material.metal(WHITE, 0.5).with_coat(1.0, 0.1),
    "fuzz over diffuse": Material.matte(WHITE, 0.0).with_fuzz(1.0, 0.5),
    "the full stack": Material.glossy(WHITE, 0.3, 0.4).with_metalness(0.3).with_coat(0.7, 0.2).with_fuzz(0.5, 0.7),
    "glass plane r=0.4": Material.glass(0.4, 1.5),
    "thin glass r=0.4": Material.glass(0.4, 1.5).thin(),
    "half opacity": Material.matte(WHITE, 0.0).with_opacity(0.5),
}


# The furnace bound per configuration, relative to the sky: 1e-3 wherever
# the tabulated energy terms allow it; the five configurations whose
# tabulated albedos leave a larger bias (measured on the oracle at 4096 spp,
# standard error below 1e-4: -1.19e-3, -1.21e-3, -1.20e-3, -1.08e-3,
# +2.61e-3) are held to that residual plus about 4e-4.
FURNACE_BOUND = {name: 1e-3 for name in FURNACE_CONFIGS}
FURNACE_BOUND.update({
    "glossy-diffuse r=0.05": 1.6e-3,
    "glossy ior=2.5": 1.6e-3,
    "coat over diffuse": 1.6e-3,
    "the full stack": 1.5e-3,
    "thin glass r=0.4": 3.0e-3,
})


def glass_sphere_scene(material: Material, sky: float = 0.5) -> SceneData:
    """render/tests.rs:499-592: a unit icosphere(3) of `material` at y = 2
    under a uniform sky, framed by a 40-degree camera 4 m away."""
    from cenote_amd.shapes import icosphere
    from cenote_amd.scene import trs

    s = SceneData()
    s.add_instance(s.add_mesh(icosphere(3)), trs((0.0, 2.0, 0.0)), material)
    s.env_texels = constant_environment((sky, sky, sky))
    s.camera = CameraData(position=(0.0, 2.0, 4.0), look_at=(0.0, 2.0, 0.0), vfov_degrees=40.0)
    return s


def color_bleed_scene() -> SceneData:
    """render/tests.rs:1051-1116: a terracotta icosphere(2) on a gray floor
    under a white sky."""
    from cenote_amd.scene import acescg_from_rec709, trs
    from cenote_amd.shapes import icosphere

    s = SceneData()
    s.add_instance(s.add_mesh(icosphere(2)), trs((0.0, 1.0, 0.0)),
                   Material.matte(acescg_from_rec709((0.7, 0.22, 0.08)), 0.6))
    s.add_instance(s.add_mesh(ground_plane(5.0)), identity(), Material.matte(acescg_from_rec709((0.65,) * 3), 0.1))
    s.env_texels = constant_environment((1.0, 1.0, 1.0))
    s.camera = CameraData(position=(0.0, 1.8, 5.0), look_at=(0.0, 1.0, 0.0), vfov_degrees=40.0)
    return s


def redness(beauty_sum, x: int, y: int) -> float:
    """render/tests.rs:1093-1104: red over blue summed over a 3x3 patch."""
    patch = beauty_sum[y:y + 3, x:x + 3]
    return float(patch[..., 0].sum() / patch[..., 2].sum())


def silhouette_scene(aperture: float = 0.0, focus: float = 0.0) -> SceneData:
    """wavefront.rs:957-1031 and :1148-1226: a gray icosphere(2) on a large
    gray floor under a constant 0.4 sky, camera 8.5 m out; with a lens of
    `aperture` radius focused at `focus` m."""
    from cenote_amd.scene import trs
    from cenote_amd.shapes import icosphere

    s = SceneData()
    s.add_instance(s.add_mesh(icosphere(2)), trs((0.0, 1.0, 0.0)), Material.matte((0.5, 0.5, 0.5), 0.3))
    s.add_instance(s.add_mesh(ground_plane(12.0)), identity(), Material.matte((0.5, 0.5, 0.5), 0.1))
    s.env_texels = constant_environment((0.4, 0.4, 0.4))
    s.camera = CameraData(position=(0.0, 2.0, 8.5), look_at=(0.0, 0.5, 0.0), vfov_degrees=40.0,
                          aperture_radius=aperture, focus_distance=focus)
    return s


SILHOUETTE_SKY = np.array([0.4, 0.4, 0.4, 1.0], np.float32)


def mixed_pixels(images) -> tuple[int, np.ndarray]:
    """Pixels that saw the sky in some samples and a surface in others
    (wavefront.rs:1011-1030), and the per-pixel saw-sky mask."""
    sky = np.stack([np.all(img == SILHOUETTE_SKY, axis=-1) for img in images])
    saw_sky, saw_surface = sky.any(axis=0), (~sky).any(axis=0)
    return int(np.sum(saw_sky & saw_surface)), sky


def _strategy_objects(s: SceneData):
    from cenote_amd.scene import trs
    from cenote_amd.shapes import icosphere

    sphere = s.add_mesh(icosphere(2))
    s.add_instance(sphere, trs((0.0, 1.0, 0.0)), Material.glossy((0.6, 0.6, 0.6), 0.4, 0.3).with_metalness(0.5))
    s.add_instance(s.add_mesh(ground_plane(4.0)), identity(), Material.glossy((0.7, 0.7, 0.7), 0.0, 0.2))
    s.camera = CameraData(position=(0.0, 2.5, 6.0), look_at=(0.0, 1.0, 0.0), vfov_degrees=45.0)
    return sphere


def strategy_scene() -> SceneData:
    """wavefront.rs:1267-1306: a half-metal sphere, a glossy floor, a glass
    ball and an emissive sphere above, under a black sky."""
    from cenote_amd.scene import trs

    s = SceneData()
    sphere = _strategy_objects(s)
    s.add_instance(sphere, trs((1.6, 0.6, 1.2), scale=(0.6, 0.6, 0.6)), Material.glass(0.4, 1.5))
    s.add_instance(sphere, trs((0.0, 3.0, 0.0), scale=(0.7, 0.7, 0.7)), Material.emitter((4.0, 4.0, 4.0)))
    s.env_texels = constant_environment((0.0, 0.0, 0.0))
    return s


def env_strategy_scene() -> SceneData:
    """wavefront.rs:1318-1356: the sphere and floor lit only by an 8x4 sky
    of 0.2 with a hard-zero upper band holding one 8.0 sun texel."""
    s = SceneData()
    _strategy_objects(s)
    tex = np.full((4, 8, 4), 0.2, np.float32)
    tex[..., 3] = 0.2
    tex[1, :, :3] = 0.0
    tex[1, 4, :3] = 8.0
    s.env_texels = tex
    return s
