"""Launch-shape autotuner for the wavefront stages.

The stage kernels' launch shape — the persistent trace grid, the
shade_surface workgroup variant (1024/512/256 lanes with the BSDF tables in
LDS, or 256 lanes reading them through the caches), the shading grid caps
the BVH width the traversal walks and the bounce from which one tail launch
finishes every live path — does not change any result: every shape renders
bitwise the same film. It does change speed, and the best shape depends on the scene (BVH
depth, material mix) and on how many paths a wave holds. ``tune`` times a
candidate grid on the caller's own workload and leaves the renderer on the
fastest shape; ``tune_capacity`` picks the path-pool (and so queue)
capacity the same way. The reference has no counterpart (its dispatch sizes come
from its queue headers, wavefront.rs:292-356).
"""

from __future__ import annotations

import itertools
import time

from . import _abi
from .render import Film, Renderer, Scene

SHADE_VARIANTS = {
    _abi.SHADE_1024_LDS: 1024,
    _abi.SHADE_512_LDS: 512,
    _abi.SHADE_256_LDS: 256,
    _abi.SHADE_256_GLOBAL: 256,
    _abi.SHADE_512_LDS_SORTED: 512,
    _abi.SHADE_256_LDS_SORTED: 256,
    _abi.SHADE_256_GLOBAL_SORTED: 256,
}


def candidate_shapes(compute_units: int, trace_per_cu=(16, 32, 64), variants=tuple(SHADE_VARIANTS), widths=(4, 8),
                     tails=(_abi.DEFAULT_TAIL_BOUNCE,)):
    """The default search grid: trace workgroups per CU x shade variants x
    BVH width (the 4- or 8-wide tree a scene carries) x tail bounce (0: the
    wavefront stages throughout; by default only the library's own), with
    each shade variant's grid capped at 2048 lanes per CU (4096 for the
    256-lane ones)."""
    for tpc, variant, width, tail in itertools.product(trace_per_cu, variants, widths, tails):
        block = SHADE_VARIANTS[variant]
        yield {
            "trace_blocks": compute_units * tpc,
            "shade_variant": variant,
            "shade_blocks": compute_units * ((4096 if block == 256 else 2048) // block),
            "light_blocks": compute_units * 4,
            "bvh_width": width,
            "tail_bounce": tail,
        }


def time_shape(renderer: Renderer, scene: Scene, film: Film, samples: int, repeats: int) -> float:
    """Best wall time of ``repeats`` stream-ordered renders of ``samples``
    samples, after one untimed render in the same shape."""
    renderer.accumulate(scene, film, count=samples)
    best = float("inf")
    for _ in range(repeats):
        film.reset()
        renderer.synchronize()
        t0 = time.perf_counter()
        renderer.accumulate(scene, film, count=samples, synchronize=False)
        renderer.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def tune(renderer: Renderer, scene: Scene, width: int, height: int, samples: int = 4, repeats: int = 2,
         candidates=None) -> dict:
    """Time every candidate shape on ``samples`` samples of a width x height
    view of ``scene`` and set the fastest on ``renderer``. Returns the chosen
    shape and the per-candidate timings (seconds)."""
    _, cus = renderer.ctx.describe()
    shapes = list(candidates if candidates is not None else candidate_shapes(cus))
    if not shapes:
        raise ValueError("no candidate launch shapes")
    film = Film(renderer.ctx, width, height)
    renderer.reserve(min(samples, renderer.samples_per_wave(width, height)) * width * height, aov=True)
    timings = []
    try:
        for shape in shapes:
            renderer.shape = shape
            timings.append((time_shape(renderer, scene, film, samples, repeats), shape))
    finally:
        film.close()
    best_time, best = min(timings, key=lambda t: t[0])
    renderer.shape = best
    return {"shape": best, "seconds": best_time,
            "timings": [{"seconds": t, **shape} for t, shape in timings]}


def tune_capacity(ctx, scene: Scene, width: int, height: int, samples: int, max_bounces: int,
                  capacities=(1 << 24, 1 << 25, 1 << 26, 1 << 27), shape=None, repeats: int = 2) -> dict:
    """The path-pool capacity — and with it every stage queue's capacity,
    which is sized from the pool — that renders ``samples`` samples of a
    width x height view of ``scene`` fastest, each candidate on a renderer
    of its own in ``shape``. Returns the capacity and the timings."""
    from .render import Renderer

    timings = []
    for capacity in capacities:
        r = Renderer(ctx, capacity=capacity, max_bounces=max_bounces)
        if shape:
            r.shape = shape
        film = Film(ctx, width, height)
        r.reserve(min(samples, r.samples_per_wave(width, height)) * width * height, aov=True)
        try:
            timings.append((time_shape(r, scene, film, samples, repeats), capacity))
        finally:
            film.close()
            r.close()
    best_time, best = min(timings, key=lambda t: t[0])
    return {"capacity": best, "seconds": best_time,
            "timings": [{"seconds": t, "capacity": c} for t, c in timings]}
