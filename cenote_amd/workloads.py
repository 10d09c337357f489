"""The benchmark and test scenes, built from seeded synthetic data.

* ``cornell()`` — Cornell-style box: five walls (2 triangles each), the two
  rotated boxes of the reference corpus (tests/scenes/cornell-box/scene-v4.pbrt
  geometry), a 0.25 x 0.25 quad light of radiance (17, 12, 4); diffuse
  albedos 0.73 gray, (0.65, 0.05, 0.05) red, (0.12, 0.45, 0.15) green; pinhole.
* ``chart()`` — material chart: 5 x 5 UV spheres (32 x 16) on a 20 x 20
  ground quad, roughness 0..1 across columns, metalness 0..1 across rows,
  base color (0.8, 0.35, 0.2); a quad key light (scene/demo.rs:124-140) and a
  seeded procedural 1024 x 512 sky with a Gaussian sun of peak 1e4; thin lens
  of aperture 0.02 focused at the look-at point.
* ``heavy()`` — 64 seeded noise-displaced geodesic spheres (32,000 triangles
  each, 2,048,000 in all) in a 50^3 box — rough glass with interior
  absorption, coat over conductor, and fuzz — plus 10,000 emissive triangles
  of log-uniform radiance in [1e-2, 1e2] and the chart's sky.
* ``furnace()`` — render/tests.rs:38-78: one big plane under a 0.5 sky.
* ``batched_views()`` — 64 seeded views of the chart and heavy scenes:
  pixel counts uniform in [128^2, 1024^2], aspect uniform in [1, 16/9], spp
  uniform in 1..64, cameras turned about each scene's look-at point.
"""

from __future__ import annotations

import numpy as np

from .scene import (CameraData, Material, MeshData, OpenPBR, SceneData, constant_environment,
                    identity, trs)
from .shapes import geodesic_sphere, ground_plane, uv_sphere, vertex_normals

# The workload shapes the benchmark and tests use (pixels, spp, depth).
# Every film accumulation writes the albedo, normal and depth guides with the
# beauty (render/film.rs:66-127), so every configuration carries AOVs.
CONFIGS = {
    "cornell": dict(width=256, height=256, spp=16, max_bounces=8, aov=True),
    "chart": dict(width=1280, height=720, spp=64, max_bounces=12, aov=True),
    "heavy": dict(width=1920, height=1080, spp=32, max_bounces=16, aov=True),
    # The 2M-triangle scene seen along the box's diagonal, so that geometry
    # covers most pixels (72% of camera rays hit, against 25% for the
    # default view): a traversal-bound view of the same scene.
    "heavy_field": dict(width=1920, height=1080, spp=32, max_bounces=16, aov=True),
}


def _quad_mesh(corners) -> MeshData:
    p = np.asarray(corners, dtype=np.float32)
    n = np.cross(p[1] - p[0], p[2] - p[0]).astype(np.float64)
    n = (n / np.linalg.norm(n)).astype(np.float32)
    return MeshData(p, np.tile(n, (4, 1)), np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32))


def _box_mesh(points, order) -> MeshData:
    """A box from the corpus' 24 corner points (4 per face) and index triples,
    with flat per-face normals facing out."""
    p = np.asarray(points, dtype=np.float32).reshape(-1, 3)
    idx = np.asarray(order, dtype=np.uint32).reshape(-1, 3)
    center = p.mean(axis=0)
    normals = np.zeros_like(p)
    for t in idx:
        a, b, c = p[t[0]], p[t[1]], p[t[2]]
        n = np.cross(b - a, c - a).astype(np.float64)
        n /= np.linalg.norm(n)
        if np.dot(n, (a + b + c) / 3 - center) < 0:
            n = -n
        normals[t] = n
    # wind every face counter-clockwise seen from outside
    fixed = []
    for t in idx:
        a, b, c = p[t[0]], p[t[1]], p[t[2]]
        n = np.cross(b - a, c - a)
        fixed.append(t if np.dot(n, (a + b + c) / 3 - center) > 0 else t[[0, 2, 1]])
    return MeshData(p, normals, np.array(fixed, dtype=np.uint32))


_SHORT_BOX = [
    -0.0460751, 0.6, 0.573007, -0.0460751, -2.98023e-08, 0.573007, 0.124253, 0, 0.00310463,
    0.124253, 0.6, 0.00310463, 0.533009, 0, 0.746079, 0.533009, 0.6, 0.746079, 0.703337, 0.6,
    0.176177, 0.703337, 2.98023e-08, 0.176177, 0.533009, 0.6, 0.746079, -0.0460751, 0.6, 0.573007,
    0.124253, 0.6, 0.00310463, 0.703337, 0.6, 0.176177, 0.703337, 2.98023e-08, 0.176177, 0.124253,
    0, 0.00310463, -0.0460751, -2.98023e-08, 0.573007, 0.533009, 0, 0.746079, 0.533009, 0,
    0.746079, -0.0460751, -2.98023e-08, 0.573007, -0.0460751, 0.6, 0.573007, 0.533009, 0.6,
    0.746079, 0.703337, 0.6, 0.176177, 0.124253, 0.6, 0.00310463, 0.124253, 0, 0.00310463,
    0.703337, 2.98023e-08, 0.176177,
]
_TALL_BOX = [
    -0.720444, 1.2, -0.473882, -0.720444, 0, -0.473882, -0.146892, 0, -0.673479, -0.146892, 1.2,
    -0.673479, -0.523986, 0, 0.0906493, -0.523986, 1.2, 0.0906492, 0.0495656, 1.2, -0.108948,
    0.0495656, 0, -0.108948, -0.523986, 1.2, 0.0906492, -0.720444, 1.2, -0.473882, -0.146892, 1.2,
    -0.673479, 0.0495656, 1.2, -0.108948, 0.0495656, 0, -0.108948, -0.146892, 0, -0.673479,
    -0.720444, 0, -0.473882, -0.523986, 0, 0.0906493, -0.523986, 0, 0.0906493, -0.720444, 0,
    -0.473882, -0.720444, 1.2, -0.473882, -0.523986, 1.2, 0.0906492, 0.0495656, 1.2, -0.108948,
    -0.146892, 1.2, -0.673479, -0.146892, 0, -0.673479, 0.0495656, 0, -0.108948,
]
_BOX_ORDER = [0, 2, 1, 0, 3, 2, 4, 6, 5, 4, 7, 6, 8, 10, 9, 8, 11, 10, 12, 14, 13, 12, 15, 14,
              16, 18, 17, 16, 19, 18, 20, 22, 21, 20, 23, 22]


def cornell() -> SceneData:
    s = SceneData()
    gray = OpenPBR(base_color=(0.73, 0.73, 0.73), specular_weight=0.0).lower()
    red = OpenPBR(base_color=(0.65, 0.05, 0.05), specular_weight=0.0).lower()
    green = OpenPBR(base_color=(0.12, 0.45, 0.15), specular_weight=0.0).lower()
    walls = [
        ([[-1, 0, -1], [-1, 0, 1], [1, 0, 1], [1, 0, -1]], gray),      # floor, normal +y
        ([[-1, 2, -1], [1, 2, -1], [1, 2, 1], [-1, 2, 1]], gray),      # ceiling, normal -y
        ([[-1, 0, -1], [1, 0, -1], [1, 2, -1], [-1, 2, -1]], gray),    # back, normal +z
        ([[-1, 0, -1], [-1, 2, -1], [-1, 2, 1], [-1, 0, 1]], red),     # left, normal +x
        ([[1, 0, -1], [1, 0, 1], [1, 2, 1], [1, 2, -1]], green),       # right, normal -x
    ]
    for corners, material in walls:
        s.add_instance(s.add_mesh(_quad_mesh(corners)), identity(), material)
    s.add_instance(s.add_mesh(_box_mesh(_SHORT_BOX, _BOX_ORDER)), identity(), gray)
    s.add_instance(s.add_mesh(_box_mesh(_TALL_BOX, _BOX_ORDER)), identity(), gray)
    h = 0.125
    light = _quad_mesh([[-h, 1.98, -h], [h, 1.98, -h], [h, 1.98, h], [-h, 1.98, h]])  # normal -y
    emitter = OpenPBR(base_color=(0.0, 0.0, 0.0), specular_weight=0.0, emission_color=(17.0, 12.0, 4.0),
                      emission_luminance=1.0).lower()
    s.add_instance(s.add_mesh(light), identity(), emitter)
    s.env_texels = constant_environment((0.0, 0.0, 0.0))
    s.camera = CameraData(position=(0.0, 1.0, 6.8), look_at=(0.0, 1.0, 0.0), vfov_degrees=19.5)
    return s


def sky(width: int = 1024, height: int = 512, seed: int = 0) -> np.ndarray:
    """A seeded procedural equirect sky (row 0 = +Y, center column = -Z,
    environment.rs:33-37): a zenith-to-horizon gradient over a darker ground
    and a Gaussian sun disc of peak radiance 1e4."""
    rng = np.random.default_rng(seed)
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = np.pi * v[:, None]
    phi = 2.0 * np.pi * (u[None, :] - 0.5)
    d = np.stack([np.sin(theta) * np.sin(phi), np.cos(theta) * np.ones_like(phi),
                  -np.sin(theta) * np.cos(phi)], axis=-1)
    up = np.clip(d[..., 1], -1.0, 1.0)
    zenith = np.array([0.18, 0.32, 0.85])
    horizon = np.array([0.75, 0.82, 0.95])
    ground = np.array([0.16, 0.14, 0.12])
    t = np.clip(up, 0.0, 1.0)[..., None] ** 0.5
    color = np.where(up[..., None] >= 0.0, horizon * (1 - t) + zenith * t,
                     ground + (horizon - ground) * np.exp(8.0 * up[..., None]))
    elevation = np.radians(rng.uniform(25.0, 50.0))
    azimuth = rng.uniform(-np.pi, np.pi)
    sun = np.array([np.cos(elevation) * np.sin(azimuth), np.sin(elevation), -np.cos(elevation) * np.cos(azimuth)])
    angle = np.arccos(np.clip(d @ sun, -1.0, 1.0))
    sigma = np.radians(0.6)
    color = color + (1e4 * np.exp(-0.5 * (angle / sigma) ** 2))[..., None] * np.array([1.0, 0.92, 0.8])
    out = np.ones((height, width, 4), dtype=np.float32)
    out[..., :3] = color.astype(np.float32)
    return out


def chart(seed: int = 0) -> SceneData:
    s = SceneData()
    sphere = s.add_mesh(uv_sphere(32, 16))
    for row in range(5):
        for col in range(5):
            material = OpenPBR(base_color=(0.8, 0.35, 0.2), specular_roughness=col / 4.0,
                               base_metalness=row / 4.0).lower()
            s.add_instance(sphere, trs((1.2 * (col - 2.0), 0.5, 1.2 * (row - 2.0)), scale=(0.5, 0.5, 0.5)),
                           material)
    floor = OpenPBR(base_color=(0.65, 0.65, 0.65), base_diffuse_roughness=0.1, specular_roughness=0.15).lower()
    s.add_instance(s.add_mesh(ground_plane(10.0)), identity(), floor)
    key = OpenPBR(base_color=(0.0, 0.0, 0.0), specular_weight=0.0, emission_color=(1.0, 0.85, 0.6),
                  emission_luminance=18.0).lower()
    s.add_instance(s.add_mesh(ground_plane(1.0)), trs((-3.5, 5.4, 1.0), (180.0, 0.0, 0.0), (0.75, 0.75, 0.75)),
                   key)
    s.env_texels = sky(1024, 512, seed)
    s.camera = CameraData(position=(0.0, 5.5, 11.0), look_at=(0.0, 0.0, 0.0), vfov_degrees=40.0,
                          aperture_radius=0.02)
    return s


def _noise_radius(p: np.ndarray, rng: np.random.Generator, amplitude: float) -> np.ndarray:
    """A smooth seeded displacement field on the unit sphere: a sum of random
    plane waves."""
    k = rng.normal(size=(8, 3)) * rng.uniform(2.0, 6.0, size=(8, 1))
    phase = rng.uniform(0.0, 2.0 * np.pi, size=8)
    weight = rng.uniform(0.3, 1.0, size=8)
    field = (np.sin(p @ k.T + phase) * weight).sum(axis=1) / weight.sum()
    return 1.0 + amplitude * field


def heavy(seed: int = 0, spheres: int = 64, frequency: int = 40, emitters: int = 10_000) -> SceneData:
    rng = np.random.default_rng(seed)
    s = SceneData()
    base_p, base_i = geodesic_sphere(frequency)
    for k in range(spheres):
        radius = rng.uniform(1.5, 3.5)
        p = base_p * _noise_radius(base_p, rng, 0.15)[:, None]
        mesh = MeshData(p.astype(np.float32), vertex_normals(p, base_i), base_i)
        center = rng.uniform(-25.0 + radius, 25.0 - radius, size=3)
        kind = k % 3
        if kind == 0:
            absorption = rng.uniform(0.0, 2.0, size=3)
            material = OpenPBR(transmission_weight=1.0, specular_ior=1.5,
                               specular_roughness=rng.uniform(0.0, 0.5),
                               transmission_color=tuple(np.exp(-absorption)), transmission_depth=1.0).lower()
        elif kind == 1:
            material = OpenPBR(base_color=tuple(rng.uniform(0.5, 1.0, size=3)), base_metalness=1.0,
                               specular_roughness=rng.uniform(0.1, 0.6), coat_weight=1.0,
                               coat_roughness=rng.uniform(0.0, 0.2)).lower()
        else:
            material = OpenPBR(base_color=tuple(rng.uniform(0.05, 0.6, size=3)), specular_weight=0.5,
                               fuzz_weight=1.0, fuzz_color=tuple(rng.uniform(0.5, 1.0, size=3)),
                               fuzz_roughness=rng.uniform(0.3, 0.8)).lower()
        s.add_instance(s.add_mesh(mesh), trs(tuple(center), scale=(radius, radius, radius)), material)
    # Emissive triangles: one single-triangle instance each (per-instance
    # emission is the only form the reference's scene model expresses).
    centers = rng.uniform(-25.0, 25.0, size=(emitters, 3))
    radiance = np.exp(rng.uniform(np.log(1e-2), np.log(1e2), size=emitters))
    for k in range(emitters):
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        b = np.cross(a, rng.normal(size=3))
        b /= np.linalg.norm(b)
        c = np.cross(a, b)
        size = 0.4
        corners = np.array([centers[k] + size * b, centers[k] + size * (-0.5 * b + 0.866 * c),
                            centers[k] + size * (-0.5 * b - 0.866 * c)], dtype=np.float32)
        n = np.cross(corners[1] - corners[0], corners[2] - corners[0]).astype(np.float64)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        mesh = MeshData(corners, np.tile(n, (3, 1)), np.array([[0, 1, 2]], dtype=np.uint32))
        s.add_instance(s.add_mesh(mesh), identity(), Material.emitter((radiance[k],) * 3))
    s.env_texels = sky(1024, 512, seed)
    s.camera = CameraData(position=(0.0, 8.0, 85.0), look_at=(0.0, 0.0, 0.0), vfov_degrees=40.0)
    return s


def furnace(material: Material, center=(0.0, 0.0, 0.0), scale: float = 1.0, aperture: float = 0.0,
            focus: float = 0.0, sky_radiance: float = 0.5) -> SceneData:
    """render/tests.rs:38-78: a plane of `material`, every camera ray lands on it."""
    s = SceneData()
    c = np.asarray(center, dtype=np.float64)
    t = np.zeros((3, 4))
    t[:, :3] = np.eye(3) * scale
    t[:, 3] = c
    s.add_instance(s.add_mesh(ground_plane(5.0)), t.astype(np.float32), material)
    s.env_texels = constant_environment((sky_radiance,) * 3)
    s.camera = CameraData(position=tuple(c + [0.0, scale, 0.0]), look_at=tuple(c + [0.0, 0.0, -scale]),
                          vfov_degrees=40.0, aperture_radius=aperture, focus_distance=focus)
    return s


def field_camera() -> CameraData:
    """The heavy scene's diagonal view: from just outside one corner of the
    50^3 box toward the opposite corner, 30 degrees."""
    return CameraData(position=(28.0, 28.0, 28.0), look_at=(-25.0, -25.0, -25.0), vfov_degrees=30.0)


def heavy_field(seed: int = 0) -> SceneData:
    s = heavy(seed)
    s.camera = field_camera()
    return s


def build(name: str, seed: int = 0) -> SceneData:
    return {"cornell": cornell, "chart": lambda: chart(seed), "heavy": lambda: heavy(seed),
            "heavy_field": lambda: heavy_field(seed)}[name]()


def batched_views(cameras: dict, seed: int = 0, count: int = 64, scenes=("chart", "heavy")) -> list:
    """Seeded batched-view specs: (scene name, width, height, spp, depth,
    camera), alternating scenes."""
    import dataclasses
    import math

    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        name = scenes[k % len(scenes)]
        pixels = rng.uniform(128.0 * 128.0, 1024.0 * 1024.0)
        aspect = rng.uniform(1.0, 16.0 / 9.0)
        width = max(1, int(round(math.sqrt(pixels * aspect))))
        height = max(1, int(round(pixels / width)))
        spp = int(rng.integers(1, 65))
        cam = cameras[name]
        angle = rng.uniform(-0.3, 0.3)
        p = np.asarray(cam.position, np.float64)
        c = np.asarray(cam.look_at, np.float64)
        v = p - c
        x = v[0] * math.cos(angle) + v[2] * math.sin(angle)
        z = -v[0] * math.sin(angle) + v[2] * math.cos(angle)
        turned = dataclasses.replace(cam, position=(float(c[0] + x), float(p[1]), float(c[2] + z)))
        out.append((name, width, height, spp, CONFIGS[name]["max_bounces"], turned))
    return out
