"""Serving: progressive render requests over resident scenes.

``Server`` wraps the library's request loop (cn_server_*): requests are
admitted while their films fit an HBM budget and interleaved into shared
waves at sample granularity — first frames on a lane of their own, bulk
samples in larger waves on two more — with up to three waves tracing at
once on the server's own streams. ``poisson_requests``
draws a seeded open-loop request stream; ``run_served`` replays it against
a server in real time and ``run_plain`` against the plain form — one
request at a time, one sample per wave from a reference-sized pool, as the
reference's Renderer::accumulate loop renders (render/mod.rs:245-259) —
rendering exactly the samples the served run needed for each request, so
both forms produce the same images and the comparison is at matched noise.
``plain_service`` times both plain forms — that reference-shaped loop and
the best one-request-at-a-time loop (the library's full pool, a request's
samples stream-ordered) — back to back, and ``fifo_schedule`` places those
service times under any arrival schedule.
"""

from __future__ import annotations

import ctypes as C
import dataclasses
import math
import time

import numpy as np

from . import _abi
from ._abi import check
from .render import Context, Film, Renderer, Scene
from .scene import CameraData


@dataclasses.dataclass
class RequestSpec:
    scene: str
    width: int
    height: int
    spp: int
    rel_mse: float
    camera: CameraData
    max_bounces: int
    arrival: float  # seconds after the stream's start


class Server:
    """The library's serving loop on one renderer."""

    def __init__(self, renderer: Renderer, hbm_budget_bytes: int, wave_paths: int = 0):
        self.renderer = renderer
        self._lib = renderer._lib
        handle = C.c_void_p()
        check(self._lib.cn_server_create(renderer.handle, hbm_budget_bytes, wave_paths, C.byref(handle)))
        self.handle = handle

    def clock(self) -> float:
        t = C.c_double()
        check(self._lib.cn_server_clock(self.handle, C.byref(t)))
        return t.value

    def submit(self, scene: Scene, camera: CameraData, width: int, height: int, spp: int, rel_mse: float = 0.0,
               max_bounces: int = 0, arrival: float = -1.0) -> int:
        desc = _abi.RequestDesc(scene=scene.handle.value, camera=camera.record(), width=width, height=height,
                                spp=spp, max_bounces=max_bounces, rel_mse=rel_mse, arrival_seconds=arrival)
        rid = C.c_uint64()
        check(self._lib.cn_server_submit(self.handle, C.byref(desc), C.byref(rid)))
        return rid.value

    def step(self) -> int:
        busy = C.c_uint32()
        check(self._lib.cn_server_step(self.handle, C.byref(busy)))
        return busy.value

    def drain(self):
        check(self._lib.cn_server_drain(self.handle))

    def status(self, rid: int) -> _abi.RequestStatus:
        st = _abi.RequestStatus()
        check(self._lib.cn_server_status(self.handle, rid, C.byref(st)))
        return st

    def stats(self) -> _abi.ServerStats:
        st = _abi.ServerStats()
        check(self._lib.cn_server_stats_get(self.handle, C.byref(st)))
        return st

    def result(self, rid: int, width: int, height: int, aov: bool = False):
        """The request's averaged beauty (and AOVs); releases its film."""
        b = np.zeros((height, width, 4), np.float32)
        if not aov:
            check(self._lib.cn_server_result(self.handle, rid, b.ctypes.data_as(C.POINTER(C.c_float)), None, None,
                                             None))
            return b
        a = np.zeros_like(b)
        n = np.zeros_like(b)
        d = np.zeros((height, width), np.float32)
        fp = lambda x: x.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
        check(self._lib.cn_server_result(self.handle, rid, fp(b), fp(a), fp(n), fp(d)))
        return b, a, n, d

    def close(self):
        if self.handle:
            check(self._lib.cn_server_destroy(self.handle))
            self.handle = None


RESOLUTIONS = ((256, 256), (1280, 720), (1920, 1080))


def orbit(camera: CameraData, rng: np.random.Generator, spread: float = 0.15) -> CameraData:
    """The scene's camera turned about its look-at point by a seeded angle."""
    p = np.asarray(camera.position, np.float64)
    c = np.asarray(camera.look_at, np.float64)
    angle = rng.uniform(-spread, spread)
    v = p - c
    x = v[0] * math.cos(angle) + v[2] * math.sin(angle)
    z = -v[0] * math.sin(angle) + v[2] * math.cos(angle)
    return dataclasses.replace(camera, position=(float(c[0] + x), float(p[1]), float(c[2] + z)))


def poisson_requests(cameras: dict, depths: dict, rate: float, count: int, seed: int = 0,
                     resolutions=RESOLUTIONS, spp_range=(1, 256), rel_mse: float = 1e-2,
                     noise_share: float = 0.5) -> list:
    """A seeded open-loop stream: exponential inter-arrivals at ``rate``
    per second; scene, resolution and target drawn uniformly; a share of
    the requests target relMSE ``rel_mse`` (capped at the top spp), the rest
    a sample count."""
    rng = np.random.default_rng(seed)
    names = sorted(cameras)
    t = 0.0
    out = []
    for _ in range(count):
        t += rng.exponential(1.0 / rate)
        name = names[rng.integers(len(names))]
        w, h = resolutions[rng.integers(len(resolutions))]
        noisy = rng.random() < noise_share
        spp = spp_range[1] if noisy else int(rng.integers(spp_range[0], spp_range[1] + 1))
        out.append(RequestSpec(name, w, h, spp, rel_mse if noisy else 0.0, orbit(cameras[name], rng), depths[name], t))
    return out


def _latencies(records):
    first = np.array([r["first_frame"] - r["arrival"] for r in records])
    done = np.array([r["done"] - r["arrival"] for r in records])
    return {
        "first_frame_p50_ms": round(float(np.percentile(first, 50)) * 1e3, 2),
        "first_frame_p99_ms": round(float(np.percentile(first, 99)) * 1e3, 2),
        "target_p50_ms": round(float(np.percentile(done, 50)) * 1e3, 2),
        "target_p99_ms": round(float(np.percentile(done, 99)) * 1e3, 2),
    }


def summarize(records) -> dict:
    paths = sum(r["width"] * r["height"] * r["samples"] for r in records)
    start = min(r["arrival"] for r in records)
    end = max(r["done"] for r in records)
    return {"requests": len(records), "paths": paths, "seconds": round(end - start, 4),
            "mpaths_per_s": round(paths / (end - start) / 1e6, 3), **_latencies(records)}


def run_served(server: Server, scenes: dict, specs: list, take_results: bool = True) -> list:
    """Replay ``specs`` in real time: each request is submitted once the
    server clock passes its arrival (stamped with that arrival); the server
    steps whenever it has work. Returns per-request records."""
    t0 = server.clock()
    pending = list(specs)
    ids = []
    open_ids = []
    busy = 0
    while pending or busy:
        now = server.clock() - t0
        while pending and pending[0].arrival <= now:
            s = pending.pop(0)
            rid = server.submit(scenes[s.scene], s.camera, s.width, s.height, s.spp, s.rel_mse, s.max_bounces,
                                arrival=t0 + s.arrival)
            ids.append((rid, s))
            open_ids.append((rid, s))
        busy = server.step()
        if not busy and pending:
            time.sleep(max(0.0, min(0.002, pending[0].arrival - (server.clock() - t0))))
        if take_results:
            still = []
            for rid, s in open_ids:
                if server.status(rid).state == _abi.REQUEST_DONE:
                    server.result(rid, s.width, s.height)
                else:
                    still.append((rid, s))
            open_ids = still
    records = []
    for rid, s in ids:
        st = server.status(rid)
        records.append({"id": rid, "scene": s.scene, "width": s.width, "height": s.height, "samples": st.samples,
                        "arrival": st.arrival_seconds - t0, "first_frame": st.first_frame_seconds - t0,
                        "done": st.done_seconds - t0})
    return records


def run_plain(ctx: Context, scenes: dict, specs: list, samples: list, capacity: int = _abi.DEFAULT_CAPACITY) -> list:
    """The plain form on the same stream: requests in arrival order, one at
    a time, each sample its own synchronous wave (cn_accumulate) from a
    ``capacity``-path pool; request k renders ``samples[k]`` samples."""
    renderers = {}
    for s in specs:
        if s.max_bounces not in renderers:
            renderers[s.max_bounces] = Renderer(ctx, capacity=capacity, max_bounces=s.max_bounces)
    films = {}
    records = []
    t0 = time.perf_counter()
    try:
        for s, n in zip(specs, samples):
            now = time.perf_counter() - t0
            if now < s.arrival:
                time.sleep(s.arrival - now)
            key = (s.width, s.height)
            if key not in films:
                films[key] = Film(ctx, s.width, s.height)
            film = films[key]
            film.reset()
            scene = scenes[s.scene]
            scene.set_camera(s.camera)
            r = renderers[s.max_bounces]
            first = None
            for k in range(n):
                r.accumulate(scene, film)
                if k == 0:
                    first = time.perf_counter() - t0
            done = time.perf_counter() - t0
            records.append({"scene": s.scene, "width": s.width, "height": s.height, "samples": n,
                            "arrival": s.arrival, "first_frame": first, "done": done})
    finally:
        for f in films.values():
            f.close()
        for r in renderers.values():
            r.close()
    return records


def plain_service(ctx: Context, scenes: dict, specs: list, samples: list, capacity: int = 0,
                  per_sample: bool = False) -> list:
    """The plain forms, requests back to back in arrival order, one at a
    time: request k renders ``samples[k]`` samples into a reset film. With
    ``per_sample`` every sample is its own synchronous wave
    (Renderer::accumulate, render/mod.rs:245-259) from a ``capacity``-path
    pool; otherwise the first sample is its own wave (the first frame) and
    the rest go out stream-ordered from the renderer's pool, synchronized
    once. Returns each request's service time and the offset of its first
    frame from the start of its service."""
    renderers = {}
    for s in specs:
        if s.max_bounces not in renderers:
            renderers[s.max_bounces] = Renderer(ctx, capacity=capacity, max_bounces=s.max_bounces)
    films = {}
    out = []
    try:
        warm = Film(ctx, 64, 64)
        for depth, r in renderers.items():  # first launches, outside the timing
            scene = next(scenes[s.scene] for s in specs if s.max_bounces == depth)
            warm.reset()
            r.accumulate(scene, warm)
        warm.close()
        if not per_sample:
            biggest = max(s.width * s.height * n for s, n in zip(specs, samples))
            for r in renderers.values():
                r.reserve(min(biggest, r.capacity), aov=True)
        for s, n in zip(specs, samples):
            key = (s.width, s.height)
            if key not in films:
                films[key] = Film(ctx, s.width, s.height)
            film = films[key]
            scene = scenes[s.scene]
            r = renderers[s.max_bounces]
            t0 = time.perf_counter()
            film.reset()
            scene.set_camera(s.camera)
            r.accumulate(scene, film)
            first = time.perf_counter() - t0
            if n > 1:
                if per_sample:
                    for _ in range(n - 1):
                        r.accumulate(scene, film)
                else:
                    r.accumulate(scene, film, count=n - 1)
            done = time.perf_counter() - t0
            out.append({"scene": s.scene, "width": s.width, "height": s.height, "samples": n,
                        "paths": s.width * s.height * n, "first_offset": first, "service": done})
    finally:
        for f in films.values():
            f.close()
        for r in renderers.values():
            r.close()
    return out


def fifo_schedule(specs: list, service: list) -> list:
    """Records of first-in first-out service under the arrival schedule of
    ``specs``: request k starts once it has arrived and request k - 1 is
    done, and takes its measured service time."""
    records = []
    free = 0.0
    for s, svc in zip(specs, service):
        start = max(s.arrival, free)
        free = start + svc["service"]
        records.append({"scene": s.scene, "width": s.width, "height": s.height, "samples": svc["samples"],
                        "arrival": s.arrival, "first_frame": start + svc["first_offset"], "done": free})
    return records
