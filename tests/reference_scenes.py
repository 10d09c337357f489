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
