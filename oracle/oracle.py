"""TEST INFRASTRUCTURE ONLY — ctypes access to the CPU oracle
(oracle/libcn_oracle.so, built by oracle/Makefile).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this module, and only as the checker. The product never does.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "libcn_oracle.so"
REPO = ORACLE_DIR.parent
DATA = REPO / "cenote_amd" / "data"

_lib = None


def build(force: bool = False) -> Path:
    args = ["make", "-C", str(ORACLE_DIR)]
    if force:
        subprocess.run(args + ["clean"], check=True, stdout=subprocess.DEVNULL)
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


def bsdf_tables() -> np.ndarray:
    a = np.fromfile(DATA / "ggx_energy.bin", dtype="<f4")
    b = np.fromfile(DATA / "ltc_sheen.bin", dtype="<f4")
    return np.ascontiguousarray(np.concatenate([a, b]), dtype=np.float32)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = C.CDLL(str(LIB_PATH))
    fp = C.POINTER(C.c_float)
    up = C.POINTER(C.c_uint32)
    vp = C.c_void_p
    u = C.c_uint32
    L.orc_sample_2d.argtypes = [u, u, u, fp]
    L.orc_sample_1d.argtypes = [u, u, u]
    L.orc_sample_1d.restype = C.c_float
    L.orc_hash_uniform.argtypes = [u, u, u, u]
    L.orc_hash_uniform.restype = C.c_float
    L.orc_crossing_uniform.argtypes = [u, u, u, u, u]
    L.orc_crossing_uniform.restype = C.c_float
    L.orc_log2_transmittance_fixed.argtypes = [C.c_float]
    L.orc_log2_transmittance_fixed.restype = C.c_int32
    L.orc_exp2_fixed.argtypes = [C.c_int32]
    L.orc_exp2_fixed.restype = C.c_float
    for name in ("orc_lowbias32", "orc_reverse_bits"):
        getattr(L, name).argtypes = [u]
        getattr(L, name).restype = u
    for name in ("orc_laine_karras", "orc_sobol"):
        getattr(L, name).argtypes = [u, u]
        getattr(L, name).restype = u
    L.orc_build_lights.argtypes = [fp, fp, up, u, vp, u, vp, C.POINTER(C.c_double)]
    L.orc_env_tables.argtypes = [u, u, fp, fp, fp, fp, C.POINTER(C.c_double)]
    L.orc_select_probability.argtypes = [C.c_double, C.c_double]
    L.orc_select_probability.restype = C.c_float
    L.orc_luminance.argtypes = [fp]
    L.orc_luminance.restype = C.c_float
    L.orc_acescg_from_rec709.argtypes = [fp, fp]
    L.orc_camera_basis.argtypes = [vp, C.c_float, fp]
    L.orc_scene_create.argtypes = [vp, fp]
    L.orc_scene_create.restype = vp
    L.orc_scene_destroy.argtypes = [vp]
    L.orc_scene_light_count.argtypes = [vp]
    L.orc_scene_light_count.restype = u
    L.orc_scene_select_prob.argtypes = [vp]
    L.orc_scene_select_prob.restype = C.c_float
    L.orc_scene_light_records.argtypes = [vp, vp]
    L.orc_trace.argtypes = [vp, u, u, u, u, u, C.c_int, fp, fp, fp, fp, C.c_int]
    L.orc_trace_rows.argtypes = [vp, u, u, u, u, u, u, u, C.c_int, fp, fp, fp, fp, C.c_int]
    L.orc_accumulate.argtypes = [vp, u, u, u, u, u, u, fp, fp, fp, fp, C.c_int]
    L.orc_accumulate_rows.argtypes = [vp, u, u, u, u, u, u, u, u, fp, C.c_int]
    L.orc_accumulate_pixels.argtypes = [vp, u, u, up, u, u, u, u, u, fp, C.c_int]
    L.orc_first_hits.argtypes = [vp, u, u, u, up, up, fp, C.c_int]
    L.orc_env_sample.argtypes = [vp, fp, u, fp]
    _lib = L
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else C.POINTER(C.c_float)()


def threads_default() -> int:
    """CENOTE_ORACLE_THREADS, else the CPUs this process may run on, capped by
    OMP_NUM_THREADS when that is set (a host's share of its cores)."""
    if "CENOTE_ORACLE_THREADS" in os.environ:
        return max(1, int(os.environ["CENOTE_ORACLE_THREADS"]))
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def sample_2d(pixel: int, sample: int, dimension: int) -> tuple[float, float]:
    out = (C.c_float * 2)()
    lib().orc_sample_2d(pixel, sample, dimension, out)
    return out[0], out[1]


def sample_1d(pixel: int, sample: int, dimension: int) -> float:
    return lib().orc_sample_1d(pixel, sample, dimension)


def build_lights(corners: np.ndarray, emission: np.ndarray, ids: np.ndarray, deltas=None):
    """lights.rs:175-224 over raw triangles: returns (records as (n, 20) u32, total power)."""
    from cenote_amd import _abi

    corners = np.ascontiguousarray(corners, dtype=np.float32).reshape(-1, 9)
    emission = np.ascontiguousarray(emission, dtype=np.float32).reshape(-1, 3)
    ids = np.ascontiguousarray(ids, dtype=np.uint32).reshape(-1, 2)
    deltas = deltas or []
    darr = (_abi.DeltaLight * max(1, len(deltas)))()
    for k, (kind, vector, color) in enumerate(deltas):
        darr[k].kind = kind
        darr[k].vector = _abi.f3(*vector)
        darr[k].color = _abi.f3(*color)
    n = corners.shape[0] + len(deltas)
    out = (_abi.LightRecord * max(1, n))()
    total = C.c_double()
    lib().orc_build_lights(_fp(corners), _fp(emission), ids.ctypes.data_as(C.POINTER(C.c_uint32)),
                           corners.shape[0], C.cast(darr, C.c_void_p), len(deltas), C.cast(out, C.c_void_p),
                           C.byref(total))
    return np.frombuffer(bytes(out), dtype=np.uint32).reshape(-1, 20)[:n].copy(), total.value


def env_tables(texels: np.ndarray):
    texels = np.ascontiguousarray(texels, dtype=np.float32)
    h, w = texels.shape[:2]
    marginal = np.zeros(h + 1, np.float32)
    conditional = np.zeros(h * (w + 1), np.float32)
    pdfs = np.zeros(h * w, np.float32)
    power = C.c_double()
    lib().orc_env_tables(w, h, _fp(texels), _fp(marginal), _fp(conditional), _fp(pdfs), C.byref(power))
    return marginal, conditional, pdfs, power.value


class OracleScene:
    """The oracle's copy of a scene (``cenote_amd.scene.SceneData``)."""

    def __init__(self, data):
        desc, keep = data.desc()
        self._tables = bsdf_tables()
        self.handle = lib().orc_scene_create(C.cast(C.pointer(desc), C.c_void_p), _fp(self._tables))
        del keep

    def __del__(self):
        if getattr(self, "handle", None):
            lib().orc_scene_destroy(self.handle)
            self.handle = None

    @property
    def select_prob(self) -> float:
        return lib().orc_scene_select_prob(self.handle)

    def light_records(self) -> np.ndarray:
        n = lib().orc_scene_light_count(self.handle)
        buf = np.zeros((max(1, n), 20), np.uint32)
        lib().orc_scene_light_records(self.handle, buf.ctypes.data_as(C.c_void_p))
        return buf[:n]

    def trace(self, width, height, sample, max_bounces=8, light_sampling=0, aov=False, threads=None):
        threads = threads or threads_default()
        rad = np.zeros((height, width, 4), np.float32)
        if not aov:
            lib().orc_trace(self.handle, width, height, sample, max_bounces, light_sampling, 0, _fp(rad),
                            None, None, None, threads)
            return rad
        alb = np.zeros((height, width, 4), np.float32)
        nrm = np.zeros((height, width, 4), np.float32)
        dep = np.zeros((height, width), np.float32)
        lib().orc_trace(self.handle, width, height, sample, max_bounces, light_sampling, 1, _fp(rad), _fp(alb),
                        _fp(nrm), _fp(dep), threads)
        return rad, alb, nrm, dep

    def trace_rows(self, width, height, row0, row1, sample, max_bounces=8, light_sampling=0, aov=False,
                   threads=None):
        threads = threads or threads_default()
        rows = row1 - row0
        rad = np.zeros((rows, width, 4), np.float32)
        alb = np.zeros((rows, width, 4), np.float32) if aov else None
        nrm = np.zeros((rows, width, 4), np.float32) if aov else None
        dep = np.zeros((rows, width), np.float32) if aov else None
        lib().orc_trace_rows(self.handle, width, height, row0, row1, sample, max_bounces, light_sampling,
                             1 if aov else 0, _fp(rad), _fp(alb), _fp(nrm), _fp(dep), threads)
        return rad

    def accumulate(self, width, height, first, count, max_bounces=8, light_sampling=0, sums=None, threads=None):
        threads = threads or threads_default()
        if sums is None:
            sums = (np.zeros((height, width, 4), np.float32), np.zeros((height, width, 4), np.float32),
                    np.zeros((height, width, 4), np.float32), np.zeros((height, width), np.float32))
        b, a, n, d = sums
        lib().orc_accumulate(self.handle, width, height, first, count, max_bounces, light_sampling, _fp(b), _fp(a),
                             _fp(n), _fp(d), threads)
        return sums

    def accumulate_rows(self, width, height, row0, row1, first, count, max_bounces=8, light_sampling=0,
                        threads=None):
        """Beauty sums of samples first..first+count-1 over rows [row0, row1)."""
        threads = threads or threads_default()
        sums = np.zeros((row1 - row0, width, 4), np.float32)
        lib().orc_accumulate_rows(self.handle, width, height, row0, row1, first, count, max_bounces,
                                  light_sampling, _fp(sums), threads)
        return sums

    def accumulate_pixels(self, width, height, pixels, first, count, max_bounces=8, light_sampling=0,
                          threads=None):
        """Beauty sums of samples first..first+count-1 at the listed pixel
        indices (n,), returned as (n, 4)."""
        threads = threads or threads_default()
        pixels = np.ascontiguousarray(pixels, dtype=np.uint32)
        sums = np.zeros((len(pixels), 4), np.float32)
        lib().orc_accumulate_pixels(self.handle, width, height, pixels.ctypes.data_as(C.POINTER(C.c_uint32)),
                                    len(pixels), first, count, max_bounces, light_sampling, _fp(sums), threads)
        return sums

    def sample_environment(self, u: np.ndarray):
        """Environment::sample for uniform pairs ``u`` (n x 2), laid out as
        Scene.sample_environment returns it."""
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 2)
        out = np.zeros((len(u), 12), dtype=np.float32)
        lib().orc_env_sample(self.handle, _fp(u), len(u), _fp(out))
        idx = out[:, 9:11].view(np.uint32)
        return out[:, 0:3], out[:, 3], out[:, 4:7], out[:, 7:9], idx[:, 0].copy(), idx[:, 1].copy()

    def first_hits(self, width, height, sample=0, threads=None):
        threads = threads or threads_default()
        inst = np.zeros((height, width), np.uint32)
        prim = np.zeros((height, width), np.uint32)
        bary = np.zeros((height, width, 2), np.float32)
        lib().orc_first_hits(self.handle, width, height, sample, inst.ctypes.data_as(C.POINTER(C.c_uint32)),
                             prim.ctypes.data_as(C.POINTER(C.c_uint32)), _fp(bary), threads)
        return inst, prim, bary
