"""Scene descriptions on the host: meshes, instances, materials, lights,
environment and camera, flattened into the records the C ABI takes.

Mirrors the procedural half of the reference's scene API
(crates/cenote/src/scene/mod.rs:58-87 ``Mesh``/``Object``,
material.rs:122-196 constructors, environment.rs:49-79, scene/mod.rs:397-447
``Camera``/``Lens``) and the authoring-side lowering of
scene/lower.rs:316-399 with the Rec.709 -> ACEScg conversion of
color.rs:13-31. Host math runs in f32 in glam's operation order where the
reference's outputs depend on it.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace

import numpy as np

from . import _abi

f32 = np.float32

# color.rs:13-17, by columns.
_ACESCG_COLS = (
    (f32(0.6130974), f32(0.07019373), f32(0.020615594)),
    (f32(0.33952314), f32(0.9163539), f32(0.10956978)),
    (f32(0.04737945), f32(0.013452399), f32(0.86981466)),
)


def acescg_from_rec709(c) -> tuple[float, float, float]:
    """color.rs:23-25 — glam's Mat3 * Vec3: x_axis*x + y_axis*y + z_axis*z in f32."""
    x, y, z = (f32(v) for v in c)
    cx, cy, cz = _ACESCG_COLS
    return tuple(float((cx[i] * x + cy[i] * y) + cz[i] * z) for i in range(3))


def luminance(c) -> float:
    """color.rs:29-31 in f32, left to right."""
    r, g, b = (f32(v) for v in c)
    return float(r * f32(0.2722287) + g * f32(0.6740818) + b * f32(0.0536895))


@dataclass
class Material:
    """The GPU record's fields (material.rs:28-109), ACEScg, already lowered."""

    base_color: tuple = (0.8, 0.8, 0.8)
    base_roughness: float = 0.0
    emission: tuple = (0.0, 0.0, 0.0)
    metalness: float = 0.0
    specular_weight: float = 0.0
    specular_roughness: float = 0.0
    specular_ior: float = 1.5
    transmission_weight: float = 0.0
    transmission_color: tuple = (1.0, 1.0, 1.0)
    transmission_depth: float = 0.0
    coat_color: tuple = (1.0, 1.0, 1.0)
    coat_weight: float = 0.0
    coat_roughness: float = 0.0
    coat_ior: float = 1.6
    coat_darkening: float = 1.0
    fuzz_weight: float = 0.0
    fuzz_color: tuple = (1.0, 1.0, 1.0)
    fuzz_roughness: float = 0.5
    opacity: float = 1.0
    thin_walled: bool = False

    # material.rs:122-196
    @staticmethod
    def matte(base_color, base_roughness=0.0) -> "Material":
        return Material(base_color=tuple(base_color), base_roughness=base_roughness)

    @staticmethod
    def glossy(base_color, base_roughness, specular_roughness) -> "Material":
        return replace(Material.matte(base_color, base_roughness), specular_weight=1.0,
                       specular_roughness=specular_roughness)

    @staticmethod
    def metal(base_color, specular_roughness) -> "Material":
        return replace(Material.matte(base_color, 0.0), metalness=1.0, specular_weight=1.0,
                       specular_roughness=specular_roughness)

    @staticmethod
    def glass(specular_roughness, specular_ior) -> "Material":
        return replace(Material.matte((1.0, 1.0, 1.0), 0.0), transmission_weight=1.0,
                       specular_weight=1.0, specular_roughness=specular_roughness,
                       specular_ior=specular_ior)

    @staticmethod
    def emitter(emission) -> "Material":
        return replace(Material.matte((0.0, 0.0, 0.0), 0.0), emission=tuple(emission))

    # material.rs:200-245: the same surface with one layer changed.
    def with_metalness(self, metalness: float) -> "Material":
        return replace(self, metalness=metalness)

    def with_coat(self, coat_weight: float, coat_roughness: float) -> "Material":
        return replace(self, coat_weight=coat_weight, coat_roughness=coat_roughness)

    def with_fuzz(self, fuzz_weight: float, fuzz_roughness: float) -> "Material":
        return replace(self, fuzz_weight=fuzz_weight, fuzz_roughness=fuzz_roughness)

    def with_ior(self, specular_ior: float) -> "Material":
        return replace(self, specular_ior=specular_ior)

    def with_opacity(self, opacity: float) -> "Material":
        return replace(self, opacity=opacity)

    def thin(self) -> "Material":
        return replace(self, thin_walled=True)

    def record(self) -> _abi.Material:
        m = _abi.Material()
        for name, _ in _abi.Material._fields_:
            if name.endswith("_texture"):
                setattr(m, name, _abi.TEXTURE_NONE)
                continue
            value = getattr(self, name)
            if isinstance(value, tuple):
                setattr(m, name, _abi.f3(*[float(v) for v in value]))
            elif name == "thin_walled":
                m.thin_walled = 1 if value else 0
            else:
                setattr(m, name, float(value))
        return m


@dataclass
class OpenPBR:
    """The authoring-side material and its defaults (scene/description.rs:250-341),
    colors in linear Rec.709."""

    base_color: tuple = (0.8, 0.8, 0.8)
    base_diffuse_roughness: float = 0.0
    base_metalness: float = 0.0
    specular_weight: float = 1.0
    specular_roughness: float = 0.3
    specular_ior: float = 1.5
    transmission_weight: float = 0.0
    transmission_color: tuple = (1.0, 1.0, 1.0)
    transmission_depth: float = 0.0
    coat_weight: float = 0.0
    coat_color: tuple = (1.0, 1.0, 1.0)
    coat_roughness: float = 0.0
    coat_ior: float = 1.6
    coat_darkening: float = 1.0
    fuzz_weight: float = 0.0
    fuzz_color: tuple = (1.0, 1.0, 1.0)
    fuzz_roughness: float = 0.5
    emission_luminance: float = 0.0
    emission_color: tuple = (1.0, 1.0, 1.0)
    geometry_opacity: float = 1.0
    geometry_thin_walled: bool = False

    def lower(self) -> Material:
        """scene/lower.rs:316-399 for constant (untextured) slots."""
        clamp01 = lambda v: float(min(max(f32(v), f32(0.0)), f32(1.0)))  # noqa: E731
        coat_weight = clamp01(self.coat_weight)
        coat_color = tuple(max(v, 0.0) for v in acescg_from_rec709(self.coat_color))
        emission_base = acescg_from_rec709(self.emission_color)
        # Vec3::ONE.lerp(coat_color, w) = 1 + (c - 1) * w  (glam's lerp)
        tint = tuple(float(f32(1.0) + (f32(c) - f32(1.0)) * f32(coat_weight)) for c in coat_color)
        emission = tuple(float((f32(e) * f32(self.emission_luminance)) * f32(t))
                         for e, t in zip(emission_base, tint))
        return Material(
            base_color=acescg_from_rec709(self.base_color),
            base_roughness=clamp01(self.base_diffuse_roughness),
            emission=emission,
            metalness=clamp01(self.base_metalness),
            specular_weight=max(self.specular_weight, 0.0),
            specular_roughness=self.specular_roughness,
            specular_ior=max(self.specular_ior, 1e-4),
            transmission_weight=clamp01(self.transmission_weight),
            transmission_color=tuple(min(max(v, 0.0), 1.0)
                                     for v in acescg_from_rec709(self.transmission_color)),
            transmission_depth=max(self.transmission_depth, 0.0),
            coat_color=coat_color,
            coat_weight=coat_weight,
            coat_roughness=clamp01(self.coat_roughness),
            coat_ior=max(self.coat_ior, 1.0),
            coat_darkening=clamp01(self.coat_darkening),
            fuzz_weight=clamp01(self.fuzz_weight),
            fuzz_color=tuple(max(v, 0.0) for v in acescg_from_rec709(self.fuzz_color)),
            fuzz_roughness=clamp01(self.fuzz_roughness),
            opacity=clamp01(self.geometry_opacity),
            thin_walled=self.geometry_thin_walled,
        )


@dataclass
class MeshData:
    positions: np.ndarray  # (V, 3) f32
    normals: np.ndarray    # (V, 3) f32
    indices: np.ndarray    # (T, 3) u32
    uvs: np.ndarray | None = None  # (V, 2) f32

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=np.float32)
        self.normals = np.ascontiguousarray(self.normals, dtype=np.float32)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32)
        if self.uvs is not None:
            self.uvs = np.ascontiguousarray(self.uvs, dtype=np.float32)

    @property
    def triangle_count(self) -> int:
        return int(self.indices.shape[0])


@dataclass
class InstanceData:
    mesh: int
    transform: np.ndarray  # (3, 4) rows of the object-to-world transform
    material: Material
    camera_visible: bool = True


@dataclass
class CameraData:
    """scene/mod.rs:397-447; aperture 0 is a pinhole; focus 0 focuses at look_at."""

    position: tuple = (0.0, 0.0, 5.0)
    look_at: tuple = (0.0, 0.0, 0.0)
    up: tuple = (0.0, 1.0, 0.0)
    vfov_degrees: float = 40.0
    aperture_radius: float = 0.0
    focus_distance: float = 0.0

    def record(self) -> _abi.Camera:
        return _abi.Camera(_abi.f3(*self.position), _abi.f3(*self.look_at), _abi.f3(*self.up),
                           self.vfov_degrees, self.aperture_radius, self.focus_distance)


@dataclass
class SceneData:
    """A whole scene as plain arrays — what ``Scene`` uploads and the oracle reads."""

    meshes: list = field(default_factory=list)
    instances: list = field(default_factory=list)
    delta_lights: list = field(default_factory=list)  # (kind, vector, color)
    env_texels: np.ndarray = None  # (H, W, 4) f32 ACEScg
    camera: CameraData = field(default_factory=CameraData)

    def add_mesh(self, mesh: MeshData) -> int:
        self.meshes.append(mesh)
        return len(self.meshes) - 1

    def add_instance(self, mesh: int, transform, material: Material, camera_visible=True) -> int:
        self.instances.append(InstanceData(mesh, np.asarray(transform, dtype=np.float32).reshape(3, 4),
                                           material, camera_visible))
        return len(self.instances) - 1

    @property
    def triangle_count(self) -> int:
        return sum(self.meshes[i.mesh].triangle_count for i in self.instances)

    def desc(self):
        """The ctypes cn_scene_desc, plus the objects that keep its arrays alive."""
        keep = []
        meshes = (_abi.Mesh * len(self.meshes))()
        for k, m in enumerate(self.meshes):
            keep.append(m)
            meshes[k].positions = m.positions.ctypes.data_as(C.POINTER(C.c_float))
            meshes[k].normals = m.normals.ctypes.data_as(C.POINTER(C.c_float))
            meshes[k].uvs = (m.uvs.ctypes.data_as(C.POINTER(C.c_float)) if m.uvs is not None
                             else C.POINTER(C.c_float)())
            meshes[k].indices = m.indices.ctypes.data_as(C.POINTER(C.c_uint32))
            meshes[k].vertex_count = m.positions.shape[0]
            meshes[k].triangle_count = m.indices.shape[0]
        instances = (_abi.Instance * len(self.instances))()
        for k, inst in enumerate(self.instances):
            instances[k].mesh = inst.mesh
            instances[k].camera_visible = 1 if inst.camera_visible else 0
            instances[k].object_to_world = (C.c_float * 12)(*[float(v) for v in inst.transform.reshape(-1)])
            instances[k].material = inst.material.record()
        deltas = (_abi.DeltaLight * max(1, len(self.delta_lights)))()
        for k, (kind, vector, color) in enumerate(self.delta_lights):
            deltas[k].kind = kind
            deltas[k].vector = _abi.f3(*vector)
            deltas[k].color = _abi.f3(*color)
        env = np.ascontiguousarray(self.env_texels, dtype=np.float32)
        keep += [meshes, instances, deltas, env]
        d = _abi.SceneDesc()
        d.meshes = meshes
        d.mesh_count = len(self.meshes)
        d.instances = instances
        d.instance_count = len(self.instances)
        d.delta_lights = deltas
        d.delta_light_count = len(self.delta_lights)
        d.environment = _abi.Environment(env.shape[1], env.shape[0], env.ctypes.data_as(C.POINTER(C.c_float)))
        d.camera = self.camera.record()
        return d, keep


def dirty_bits(old: SceneData, new: SceneData) -> int:
    """The ``_abi.DIRTY_*`` bits that carry a resident ``old`` scene to ``new``
    (what changeset.rs:351-356's Dirty set names, at this ABI's grain):
    geometry when a mesh, an instance's mesh or its transform differs;
    materials when a material or camera visibility differs; lights, the
    environment and the camera when they differ."""
    bits = 0
    same_meshes = len(old.meshes) == len(new.meshes) and all(
        np.array_equal(a.positions, b.positions) and np.array_equal(a.normals, b.normals)
        and np.array_equal(a.indices, b.indices) for a, b in zip(old.meshes, new.meshes))
    same_placement = len(old.instances) == len(new.instances) and all(
        a.mesh == b.mesh and np.array_equal(a.transform, b.transform) for a, b in zip(old.instances, new.instances))
    if not (same_meshes and same_placement):
        bits |= _abi.DIRTY_GEOMETRY
    elif any(a.material != b.material or a.camera_visible != b.camera_visible
             for a, b in zip(old.instances, new.instances)):
        bits |= _abi.DIRTY_MATERIALS
    if [(k, tuple(v), tuple(c)) for k, v, c in old.delta_lights] != \
            [(k, tuple(v), tuple(c)) for k, v, c in new.delta_lights]:
        bits |= _abi.DIRTY_LIGHTS
    if old.env_texels is None or new.env_texels is None or old.env_texels.shape != new.env_texels.shape or \
            not np.array_equal(old.env_texels, new.env_texels):
        bits |= _abi.DIRTY_ENVIRONMENT
    if old.camera != new.camera:
        bits |= _abi.DIRTY_CAMERA
    return bits


def constant_environment(radiance) -> np.ndarray:
    """environment.rs:54-56: a 1x1 image."""
    return np.array([[[radiance[0], radiance[1], radiance[2], 1.0]]], dtype=np.float32)


def identity() -> np.ndarray:
    return np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]], dtype=np.float32)


def trs(translate=(0.0, 0.0, 0.0), rotate_degrees=(0.0, 0.0, 0.0), scale=(1.0, 1.0, 1.0)) -> np.ndarray:
    """scene/description.rs:137-150: T * Rz * Ry * Rx * S, rows of the 3x4 (computed in f64)."""
    rx, ry, rz = np.radians(np.asarray(rotate_degrees, dtype=np.float64))
    cx, sx, cy, sy, cz, sz = np.cos(rx), np.sin(rx), np.cos(ry), np.sin(ry), np.cos(rz), np.sin(rz)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    m = Rz @ Ry @ Rx @ np.diag(np.asarray(scale, dtype=np.float64))
    out = np.zeros((3, 4))
    out[:, :3] = m
    out[:, 3] = translate
    return out.astype(np.float32)
