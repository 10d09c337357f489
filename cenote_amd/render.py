"""The host API over libcenote_hip.so, shaped like the reference crate's:
``Context`` (gpu/mod.rs:93), ``Scene`` (scene/mod.rs:254), ``Wavefront`` /
``Renderer`` (wavefront.rs:486-629, render/mod.rs:136-299), ``Film``
(render/film.rs:66-210) and ``LightSampling`` (wavefront.rs:428-439).
Images are numpy arrays: (H, W, 4) RGBA f32 and (H, W) depth.
"""

from __future__ import annotations

import ctypes as C
import enum

import numpy as np

from . import _abi
from ._abi import check, library
from .scene import CameraData, SceneData


class LightSampling(enum.IntEnum):
    MIS = _abi.LIGHT_SAMPLING_MIS
    BSDF_ONLY = _abi.LIGHT_SAMPLING_BSDF
    NEE_ONLY = _abi.LIGHT_SAMPLING_NEE


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else C.POINTER(C.c_float)()


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Context:
    """One GPU (gpu/mod.rs:48-222)."""

    def __init__(self, device: int = 0):
        self._lib = library()
        handle = C.c_void_p()
        check(self._lib.cn_context_create(device, C.byref(handle)))
        self.handle = handle

    def describe(self) -> tuple[str, int]:
        name = C.create_string_buffer(128)
        cus = C.c_uint32()
        check(self._lib.cn_context_describe(self.handle, name, 128, C.byref(cus)))
        return name.value.decode(), cus.value

    def close(self):
        if self.handle:
            check(self._lib.cn_context_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sample_sequence(self, pixel: int, dimension: int, count: int):
        """The rng_test kernel (shaders/rng_test.slang): sample_2d points and
        sample_1d values of one key, computed on the device."""
        points = np.zeros((count, 2), dtype=np.float32)
        values = np.zeros(count, dtype=np.float32)
        check(self._lib.cn_sample_sequence(self.handle, pixel, dimension, count, _fp(points), _fp(values)))
        return points, values


class Scene:
    """A scene resident on the device (scene/mod.rs:161-219)."""

    def __init__(self, ctx: Context, data: SceneData, bvh_builder: int = _abi.BVH_DEVICE):
        """``bvh_builder``: ``_abi.BVH_DEVICE`` (default) or ``_abi.BVH_HOST``;
        both build the same tree."""
        self.ctx = ctx
        self._lib = ctx._lib
        desc, keep = data.desc()
        desc.bvh_builder = bvh_builder
        handle = C.c_void_p()
        check(self._lib.cn_scene_create(ctx.handle, C.byref(desc), C.byref(handle)))
        del keep
        self.handle = handle

    def set_camera(self, camera: CameraData):
        cam = camera.record()
        check(self._lib.cn_scene_set_camera(self.handle, C.byref(cam)))

    def update(self, data: SceneData, dirty: int):
        """Scene::update (scene/prep.rs:122-177): rebuild what ``dirty``
        (``_abi.DIRTY_*`` bits) names from the edited ``data``; the rest stays
        resident. A rejected edit (CenoteError with CN_E_SCENE) leaves the
        scene as it was."""
        desc, keep = data.desc()
        check(self._lib.cn_scene_update(self.handle, C.byref(desc), dirty))
        del keep

    def bvh(self):
        """The resident BVH: nodes as (N, 32) u32 words (128-byte records)
        and leaf-order triangles as (M, 12) u32 words (48-byte records)."""
        nn, nt = C.c_uint32(), C.c_uint32()
        check(self._lib.cn_scene_bvh(self.handle, None, 0, C.byref(nn), None, 0, C.byref(nt)))
        nodes = np.zeros((nn.value, 32), np.uint32)
        tris = np.zeros((nt.value, 12), np.uint32)
        check(self._lib.cn_scene_bvh(self.handle, nodes.ctypes.data, nn.value, C.byref(nn), tris.ctypes.data,
                                     nt.value, C.byref(nt)))
        return nodes, tris

    def bvh8(self):
        """The 8-wide traversal nodes as (N, 32) u32 words (128-byte
        BvhNode8Q records; N = 0 when the scene has no 8-wide tree), over
        the same leaf-order triangles as ``bvh()``."""
        nn = C.c_uint32()
        check(self._lib.cn_scene_bvh8(self.handle, None, 0, C.byref(nn)))
        nodes = np.zeros((nn.value, 32), np.uint32)
        check(self._lib.cn_scene_bvh8(self.handle, nodes.ctypes.data, nn.value, C.byref(nn)))
        return nodes

    def info(self) -> _abi.SceneInfo:
        info = _abi.SceneInfo()
        check(self._lib.cn_scene_info_get(self.handle, C.byref(info)))
        return info

    def light_records(self) -> np.ndarray:
        """The alias-table records as a structured u32 view (80 bytes each)."""
        count = C.c_uint32()
        check(self._lib.cn_scene_light_records(self.handle, None, 0, C.byref(count)))
        recs = (_abi.LightRecord * max(1, count.value))()
        check(self._lib.cn_scene_light_records(self.handle, recs, count.value, C.byref(count)))
        return np.frombuffer(bytes(recs), dtype=np.uint32).reshape(-1, 20)[: count.value].copy()

    def env_tables(self):
        info = self.info()
        w, h = info.env_width, info.env_height
        marginal = np.zeros(h + 1, dtype=np.float32)
        conditional = np.zeros(h * (w + 1), dtype=np.float32)
        pdfs = np.zeros(h * w, dtype=np.float32)
        check(self._lib.cn_scene_env_tables(self.handle, _fp(marginal), _fp(conditional), _fp(pdfs)))
        return marginal, conditional, pdfs

    def sample_environment(self, u: np.ndarray):
        """The device's environment draw for uniform pairs ``u`` (n x 2):
        (direction n x 3, pdf n, radiance n x 3, position n x 2 as (v, u),
        row n, column n)."""
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 2)
        out = np.zeros((len(u), 12), dtype=np.float32)
        check(self._lib.cn_scene_sample_environment(self.handle, _fp(u), len(u), _fp(out)))
        idx = out[:, 9:11].view(np.uint32)
        return out[:, 0:3], out[:, 3], out[:, 4:7], out[:, 7:9], idx[:, 0].copy(), idx[:, 1].copy()

    def close(self):
        if self.handle:
            check(self._lib.cn_scene_destroy(self.handle))
            self.handle = None


class Film:
    """Running per-pixel sums of beauty and the three AOVs (render/film.rs)."""

    def __init__(self, ctx: Context, width: int, height: int):
        self.ctx = ctx
        self._lib = ctx._lib
        self.width, self.height = width, height
        handle = C.c_void_p()
        check(self._lib.cn_film_create(ctx.handle, width, height, C.byref(handle)))
        self.handle = handle

    @property
    def samples(self) -> int:
        s = C.c_uint32()
        check(self._lib.cn_film_samples(self.handle, C.byref(s)))
        return s.value

    def reset(self):
        check(self._lib.cn_film_reset(self.handle))

    def _buffers(self):
        h, w = self.height, self.width
        return (np.zeros((h, w, 4), np.float32), np.zeros((h, w, 4), np.float32),
                np.zeros((h, w, 4), np.float32), np.zeros((h, w), np.float32))

    def sums(self):
        b, a, n, d = self._buffers()
        check(self._lib.cn_film_sums(self.handle, _fp(b), _fp(a), _fp(n), _fp(d)))
        return b, a, n, d

    def averages(self):
        """Film::averages (render/film.rs:183-197)."""
        b, a, n, d = self._buffers()
        check(self._lib.cn_film_averages(self.handle, _fp(b), _fp(a), _fp(n), _fp(d)))
        return b, a, n, d

    def set_sample(self, beauty=None, albedo=None, normal=None, depth=None):
        conv = lambda x: None if x is None else np.ascontiguousarray(x, dtype=np.float32)  # noqa: E731
        b, a, n, d = conv(beauty), conv(albedo), conv(normal), conv(depth)
        check(self._lib.cn_film_set_sample(self.handle, _fp(b), _fp(a), _fp(n), _fp(d)))

    def close(self):
        if self.handle:
            check(self._lib.cn_film_destroy(self.handle))
            self.handle = None


class Renderer:
    """The wavefront engine plus the film kernels (Wavefront::new +
    Renderer::with_max_bounces)."""

    def __init__(self, ctx: Context, capacity: int = 0,
                 max_bounces: int = _abi.DEFAULT_MAX_BOUNCES, light_sampling: LightSampling = LightSampling.MIS):
        """``capacity`` paths in the pool; 0 selects the library's pool
        (``_abi.POOL_CAPACITY``), a multiple of the reference's 2^20."""
        self.ctx = ctx
        self._lib = ctx._lib
        self.max_bounces = max_bounces
        self.capacity = capacity or _abi.POOL_CAPACITY
        handle = C.c_void_p()
        check(self._lib.cn_renderer_create(ctx.handle, capacity, max_bounces, int(light_sampling), C.byref(handle)))
        self.handle = handle

    def trace(self, scene: Scene, width: int, height: int, sample: int, aov: bool = False):
        """Wavefront::trace: one sample into a fresh radiance image; with
        ``aov`` also the wave's albedo, normal and depth accumulators."""
        rad = np.zeros((height, width, 4), np.float32)
        if not aov:
            check(self._lib.cn_trace(self.handle, scene.handle, _fp(rad), width, height, sample))
            return rad
        alb = np.zeros((height, width, 4), np.float32)
        nrm = np.zeros((height, width, 4), np.float32)
        dep = np.zeros((height, width), np.float32)
        check(self._lib.cn_trace_aov(self.handle, scene.handle, _fp(rad), _fp(alb), _fp(nrm), _fp(dep),
                                     width, height, sample))
        return rad, alb, nrm, dep

    def render(self, scene: Scene, width: int, height: int):
        """Renderer::render (render/mod.rs:209-232): one sample, sample index 0."""
        return self.trace(scene, width, height, 0)

    def first_hits(self, scene: Scene, width: int, height: int, sample: int = 0):
        inst = np.zeros((height, width), np.uint32)
        prim = np.zeros((height, width), np.uint32)
        bary = np.zeros((height, width, 2), np.float32)
        check(self._lib.cn_first_hits(self.handle, scene.handle, width, height, sample, _up(inst), _up(prim),
                                      _fp(bary)))
        return inst, prim, bary

    def accumulate(self, scene: Scene, film: Film, count: int = 1, synchronize: bool = True):
        """Renderer::accumulate, `count` times; stream-ordered, one sync at the end."""
        if count == 1 and synchronize:
            check(self._lib.cn_accumulate(self.handle, scene.handle, film.handle))
            return
        check(self._lib.cn_accumulate_async(self.handle, scene.handle, film.handle, count))
        if synchronize:
            self.synchronize()

    def render_batch(self, views):
        """Many views in shared waves: ``views`` is a sequence of
        ``(scene, film, camera, samples)`` or ``(scene, film, camera,
        samples, max_bounces)``; each film gets ``samples`` more samples of
        its view, bitwise as sequential ``accumulate`` calls on a renderer of
        that depth."""
        jobs = (_abi.ViewJob * max(1, len(views)))()
        for j, view in zip(jobs, views):
            scene, film, camera, samples = view[:4]
            j.scene = scene.handle.value
            j.film = film.handle.value
            j.camera = camera.record()
            j.samples = samples
            j.max_bounces = view[4] if len(view) > 4 else 0
        check(self._lib.cn_render_batch(self.handle, jobs, len(views)))

    def accumulate_sample(self, film: Film):
        check(self._lib.cn_film_accumulate_sample(self.handle, film.handle))

    def synchronize(self):
        check(self._lib.cn_renderer_synchronize(self.handle))

    def resolve(self, film: Film):
        b, a, n, d = film._buffers()
        check(self._lib.cn_resolve(self.handle, film.handle, _fp(b), _fp(a), _fp(n), _fp(d)))
        return b, a, n, d

    def reserve(self, paths: int, aov: bool = True):
        """Pre-size the pool and slot targets for waves of up to ``paths``."""
        check(self._lib.cn_renderer_reserve(self.handle, paths, 1 if aov else 0))

    def samples_per_wave(self, width: int, height: int) -> int:
        """How many samples of a width x height view one wave carries."""
        return max(1, self.capacity // (width * height))

    @property
    def shape(self) -> dict:
        """The launch shape: trace_blocks, shade_variant, shade_blocks, light_blocks, bvh_width,
        tail_bounce."""
        sh = _abi.LaunchShape()
        check(self._lib.cn_renderer_get_shape(self.handle, C.byref(sh)))
        return {name: getattr(sh, name) for name, _ in _abi.LaunchShape._fields_}

    @shape.setter
    def shape(self, value: dict):
        sh = _abi.LaunchShape(**{**self.shape, **value})
        check(self._lib.cn_renderer_set_shape(self.handle, C.byref(sh)))

    def set_graphs(self, enabled: bool):
        """Replay repeated wave shapes from captured hipGraphs."""
        check(self._lib.cn_renderer_set_graphs(self.handle, 1 if enabled else 0))

    def graph_stats(self):
        captures, replays = C.c_uint64(), C.c_uint64()
        check(self._lib.cn_renderer_graph_stats(self.handle, C.byref(captures), C.byref(replays)))
        return captures.value, replays.value

    def set_counters(self, enabled: bool):
        check(self._lib.cn_renderer_set_counters(self.handle, 1 if enabled else 0))

    def wave_counts(self) -> dict:
        wc = _abi.WaveCounts()
        check(self._lib.cn_renderer_wave_counts(self.handle, C.byref(wc)))
        b = wc.bounces
        return {k: np.array(getattr(wc, k)[:b], dtype=np.int64) for k in ("rays", "hits", "misses", "shadows")}

    def set_profiling(self, enabled: bool):
        check(self._lib.cn_renderer_set_profiling(self.handle, 1 if enabled else 0))

    def kernel_times(self) -> dict:
        """{kernel: (total_ms, launches)} since the last call."""
        kt = _abi.KernelTimes()
        check(self._lib.cn_renderer_kernel_times(self.handle, C.byref(kt)))
        return {bytes(kt.names[k]).split(b"\0")[0].decode(): (kt.total_ms[k], int(kt.launches[k]))
                for k in range(kt.kernels)}

    def traversal_counts(self) -> dict:
        tc = _abi.TraversalCounts()
        check(self._lib.cn_renderer_traversal_counts(self.handle, C.byref(tc)))
        return {k: int(getattr(tc, k)) for k, _ in _abi.TraversalCounts._fields_}

    def traversal_counts_by_bounce(self) -> np.ndarray:
        """(max_bounces, 4) int64: per bounce, ray nodes, ray triangles,
        shadow nodes, shadow triangles (bounce b's shadow columns count the
        connections bounce b - 1 made)."""
        out = np.zeros(4 * self.max_bounces, dtype=np.uint64)
        check(self._lib.cn_renderer_traversal_counts_by_bounce(
            self.handle, out.ctypes.data_as(C.POINTER(C.c_uint64)), self.max_bounces))
        return out.astype(np.int64).reshape(self.max_bounces, 4)

    def close(self):
        if self.handle:
            check(self._lib.cn_renderer_destroy(self.handle))
            self.handle = None
