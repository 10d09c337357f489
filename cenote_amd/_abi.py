"""ctypes mirror of include/cenote_hip.h and the loader for libcenote_hip.so.

The library is built in-tree (``cenote_amd/libcenote_hip.so``) by
``cenote_amd.build``. Loading fails loudly when it is missing: there is no
CPU fallback on this path.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE_DIR / "libcenote_hip.so"

CN_OK = 0
CN_E_NO_DEVICE = 1
CN_E_ALLOC = 2
CN_E_SCENE = 3
CN_E_INVALID = 4
CN_E_DEVICE_FAULT = 5

LIGHT_SAMPLING_MIS = 0
LIGHT_SAMPLING_BSDF = 1
LIGHT_SAMPLING_NEE = 2

LIGHT_DISTANT = 1
LIGHT_POINT = 2

TEXTURE_NONE = 0xFFFFFFFF
BVH_DEVICE = 0
BVH_HOST = 1
DIRTY_CAMERA = 1
DIRTY_MATERIALS = 2
DIRTY_LIGHTS = 4
DIRTY_ENVIRONMENT = 8
DIRTY_GEOMETRY = 16
DIRTY_ALL = 31
DEFAULT_CAPACITY = 1 << 20
DEFAULT_MAX_BOUNCES = 8
POOL_CAPACITY = 1 << 27
DEFAULT_TAIL_BOUNCE = 6
STATIC_MAX_PIXELS = 1 << 18

f3 = C.c_float * 3


class Material(C.Structure):
    """The 144-byte lowered OpenPBR record (crates/cenote/src/material.rs:28-116)."""

    _fields_ = [
        ("base_color", f3),
        ("base_roughness", C.c_float),
        ("emission", f3),
        ("metalness", C.c_float),
        ("specular_weight", C.c_float),
        ("specular_roughness", C.c_float),
        ("specular_ior", C.c_float),
        ("transmission_weight", C.c_float),
        ("transmission_color", f3),
        ("transmission_depth", C.c_float),
        ("coat_color", f3),
        ("coat_weight", C.c_float),
        ("coat_roughness", C.c_float),
        ("coat_ior", C.c_float),
        ("coat_darkening", C.c_float),
        ("fuzz_weight", C.c_float),
        ("fuzz_color", f3),
        ("fuzz_roughness", C.c_float),
        ("opacity", C.c_float),
        ("thin_walled", C.c_uint32),
        ("base_color_texture", C.c_uint32),
        ("specular_roughness_texture", C.c_uint32),
        ("metalness_texture", C.c_uint32),
        ("emission_texture", C.c_uint32),
        ("opacity_texture", C.c_uint32),
        ("normal_texture", C.c_uint32),
    ]


class Mesh(C.Structure):
    _fields_ = [
        ("positions", C.POINTER(C.c_float)),
        ("normals", C.POINTER(C.c_float)),
        ("uvs", C.POINTER(C.c_float)),
        ("indices", C.POINTER(C.c_uint32)),
        ("vertex_count", C.c_uint32),
        ("triangle_count", C.c_uint32),
    ]


class Instance(C.Structure):
    _fields_ = [
        ("mesh", C.c_uint32),
        ("camera_visible", C.c_uint32),
        ("object_to_world", C.c_float * 12),
        ("material", Material),
    ]


class DeltaLight(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("vector", f3), ("color", f3)]


class Environment(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("texels", C.POINTER(C.c_float))]


class Camera(C.Structure):
    _fields_ = [
        ("position", f3),
        ("look_at", f3),
        ("up", f3),
        ("vfov_degrees", C.c_float),
        ("aperture_radius", C.c_float),
        ("focus_distance", C.c_float),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("meshes", C.POINTER(Mesh)),
        ("mesh_count", C.c_uint32),
        ("instances", C.POINTER(Instance)),
        ("instance_count", C.c_uint32),
        ("delta_lights", C.POINTER(DeltaLight)),
        ("delta_light_count", C.c_uint32),
        ("environment", Environment),
        ("camera", Camera),
        ("bvh_builder", C.c_uint32),
        ("pad", C.c_uint32),
    ]


class LightRecord(C.Structure):
    _fields_ = [
        ("anchor", f3),
        ("pdf", C.c_float),
        ("edge1", f3),
        ("alias_threshold", C.c_float),
        ("edge2", f3),
        ("alias_index", C.c_uint32),
        ("emission", f3),
        ("instance", C.c_uint32),
        ("kind", C.c_uint32),
        ("primitive", C.c_uint32),
        ("pad", C.c_uint32 * 2),
    ]


class SceneInfo(C.Structure):
    _fields_ = [
        ("instance_count", C.c_uint32),
        ("triangle_count", C.c_uint32),
        ("light_count", C.c_uint32),
        ("env_width", C.c_uint32),
        ("env_height", C.c_uint32),
        ("env_select_prob", C.c_float),
        ("env_power", C.c_double),
        ("bvh_node_count", C.c_uint32),
        ("bvh_width", C.c_uint32),
        ("bvh_node_bytes", C.c_uint32),
        ("bvh_triangle_bytes", C.c_uint32),
        ("bvh_max_depth", C.c_uint32),
        ("bvh_builder", C.c_uint32),
        ("bvh_build_seconds", C.c_double),
        ("prep_seconds", C.c_double),
        ("bvh8_node_count", C.c_uint32),
        ("bvh8_max_depth", C.c_uint32),
    ]


class WaveCounts(C.Structure):
    _fields_ = [
        ("bounces", C.c_uint32),
        ("rays", C.c_uint32 * 256),
        ("hits", C.c_uint32 * 256),
        ("misses", C.c_uint32 * 256),
        ("shadows", C.c_uint32 * 256),
    ]


class TraversalCounts(C.Structure):
    _fields_ = [
        ("ray_nodes", C.c_uint64),
        ("ray_triangles", C.c_uint64),
        ("shadow_nodes", C.c_uint64),
        ("shadow_triangles", C.c_uint64),
    ]


SHADE_1024_LDS = 0
SHADE_512_LDS = 1
SHADE_256_LDS = 2
SHADE_256_GLOBAL = 3
SHADE_512_LDS_SORTED = 4
SHADE_256_LDS_SORTED = 5
SHADE_256_GLOBAL_SORTED = 6


class LaunchShape(C.Structure):
    _fields_ = [
        ("trace_blocks", C.c_uint32),
        ("shade_variant", C.c_uint32),
        ("shade_blocks", C.c_uint32),
        ("light_blocks", C.c_uint32),
        ("bvh_width", C.c_uint32),
        ("tail_bounce", C.c_uint32),
    ]


class ViewJob(C.Structure):
    _fields_ = [
        ("scene", C.c_void_p),
        ("film", C.c_void_p),
        ("camera", Camera),
        ("samples", C.c_uint32),
        ("max_bounces", C.c_uint32),
    ]


REQUEST_QUEUED = 0
REQUEST_ACTIVE = 1
REQUEST_DONE = 2
REQUEST_RELEASED = 3


class RequestDesc(C.Structure):
    _fields_ = [
        ("scene", C.c_void_p),
        ("camera", Camera),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("spp", C.c_uint32),
        ("max_bounces", C.c_uint32),
        ("rel_mse", C.c_float),
        ("pad", C.c_float),
        ("arrival_seconds", C.c_double),
    ]


class RequestStatus(C.Structure):
    _fields_ = [
        ("id", C.c_uint64),
        ("state", C.c_uint32),
        ("samples", C.c_uint32),
        ("arrival_seconds", C.c_double),
        ("admitted_seconds", C.c_double),
        ("first_frame_seconds", C.c_double),
        ("done_seconds", C.c_double),
        ("rel_mse_estimate", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("waves", C.c_uint64),
        ("paths", C.c_uint64),
        ("queued", C.c_uint32),
        ("active", C.c_uint32),
        ("hbm_used_bytes", C.c_uint64),
    ]


class KernelTimes(C.Structure):
    _fields_ = [
        ("kernels", C.c_uint32),
        ("names", (C.c_char * 24) * 8),
        ("total_ms", C.c_double * 8),
        ("launches", C.c_uint64 * 8),
    ]


assert C.sizeof(Material) == 144
assert C.sizeof(LightRecord) == 80

_P = C.c_void_p
_FP = C.POINTER(C.c_float)
_UP = C.POINTER(C.c_uint32)

# (name, argtypes) for every entry point the header declares.
SIGNATURES = {
    "cn_last_error": [],
    "cn_context_create": [C.c_int, C.POINTER(_P)],
    "cn_context_destroy": [_P],
    "cn_context_describe": [_P, C.c_char_p, C.c_size_t, _UP],
    "cn_scene_create": [_P, C.POINTER(SceneDesc), C.POINTER(_P)],
    "cn_scene_destroy": [_P],
    "cn_scene_set_camera": [_P, C.POINTER(Camera)],
    "cn_scene_update": [_P, C.POINTER(SceneDesc), C.c_uint32],
    "cn_scene_info_get": [_P, C.POINTER(SceneInfo)],
    "cn_scene_bvh": [_P, _P, C.c_uint32, _UP, _P, C.c_uint32, _UP],
    "cn_scene_light_records": [_P, C.POINTER(LightRecord), C.c_uint32, _UP],
    "cn_scene_env_tables": [_P, _FP, _FP, _FP],
    "cn_bvh_quantize": [C.c_void_p, C.c_uint32, C.c_void_p],
    "cn_scene_bvh8": [_P, _P, C.c_uint32, _UP],
    "cn_scene_sample_environment": [_P, _FP, C.c_uint32, _FP],
    "cn_renderer_create": [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(_P)],
    "cn_renderer_destroy": [_P],
    "cn_renderer_reserve": [_P, C.c_uint32, C.c_int],
    "cn_renderer_get_shape": [_P, C.POINTER(LaunchShape)],
    "cn_renderer_set_shape": [_P, C.POINTER(LaunchShape)],
    "cn_renderer_set_graphs": [_P, C.c_int],
    "cn_renderer_graph_stats": [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
    "cn_renderer_set_counters": [_P, C.c_int],
    "cn_renderer_wave_counts": [_P, C.POINTER(WaveCounts)],
    "cn_renderer_traversal_counts": [_P, C.POINTER(TraversalCounts)],
    "cn_renderer_traversal_counts_by_bounce": [_P, C.POINTER(C.c_uint64), C.c_uint32],
    "cn_renderer_set_profiling": [_P, C.c_int],
    "cn_renderer_kernel_times": [_P, C.POINTER(KernelTimes)],
    "cn_calibrate_valu": [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_uint64)],
    "cn_trace": [_P, _P, _FP, C.c_uint32, C.c_uint32, C.c_uint32],
    "cn_trace_aov": [_P, _P, _FP, _FP, _FP, _FP, C.c_uint32, C.c_uint32, C.c_uint32],
    "cn_first_hits": [_P, _P, C.c_uint32, C.c_uint32, C.c_uint32, _UP, _UP, _FP],
    "cn_film_create": [_P, C.c_uint32, C.c_uint32, C.POINTER(_P)],
    "cn_film_destroy": [_P],
    "cn_film_reset": [_P],
    "cn_film_samples": [_P, _UP],
    "cn_accumulate": [_P, _P, _P],
    "cn_accumulate_async": [_P, _P, _P, C.c_uint32],
    "cn_accumulate_async_static": [_P, _P, _P, C.c_uint32],
    "cn_render_batch": [_P, C.POINTER(ViewJob), C.c_uint32],
    "cn_server_create": [_P, C.c_uint64, C.c_uint32, C.POINTER(_P)],
    "cn_server_destroy": [_P],
    "cn_server_clock": [_P, C.POINTER(C.c_double)],
    "cn_server_submit": [_P, C.POINTER(RequestDesc), C.POINTER(C.c_uint64)],
    "cn_server_step": [_P, _UP],
    "cn_server_drain": [_P],
    "cn_server_status": [_P, C.c_uint64, C.POINTER(RequestStatus)],
    "cn_server_stats_get": [_P, C.POINTER(ServerStats)],
    "cn_server_result": [_P, C.c_uint64, _FP, _FP, _FP, _FP],
    "cn_write_exr": [C.c_char_p, C.c_uint32, C.c_uint32, _FP],
    "cn_write_aov_exr": [C.c_char_p, C.c_uint32, C.c_uint32, _FP, _FP, _FP, _FP],
    "cn_read_exr": [C.c_char_p, _UP, _UP, _FP, C.c_size_t],
    "cn_renderer_synchronize": [_P],
    "cn_film_averages": [_P, _FP, _FP, _FP, _FP],
    "cn_film_sums": [_P, _FP, _FP, _FP, _FP],
    "cn_film_set_sample": [_P, _FP, _FP, _FP, _FP],
    "cn_film_accumulate_sample": [_P, _P],
    "cn_resolve": [_P, _P, _FP, _FP, _FP, _FP],
    "cn_sample_sequence": [_P, C.c_uint32, C.c_uint32, C.c_uint32, _FP, _FP],
}


class CenoteError(RuntimeError):
    """A non-CN_OK status from the library, with its message."""

    def __init__(self, status: int, message: str):
        super().__init__(f"cenote_hip status {status}: {message}")
        self.status = status


_lib = None


def library() -> C.CDLL:
    """Load libcenote_hip.so (once). Raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("CENOTE_HIP_LIBRARY", LIBRARY_PATH))
    if not path.exists():
        raise RuntimeError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(there is no CPU fallback)"
        )
    lib = C.CDLL(str(path))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name == "cn_last_error" else C.c_int
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != CN_OK:
        raise CenoteError(status, library().cn_last_error().decode(errors="replace"))
