"""Benchmark of the wavefront path-tracing hot path on one MI355X.

A step is one wave: one camera path per pixel of the workload's target,
traced to the depth cap and folded into the film (Renderer::accumulate,
render/mod.rs:245-259). The default workload is the material chart —
1280x720, depth 12, thin lens 0.02, seeded procedural sky — and the default
run times 64 steps, one full 64-spp render. The JSON line reports camera
paths per second over the timed region, the dominant kernel's algorithmic
bandwidth against the HBM roofline (HIP events on the launch stream), and
the CPU oracle timed on a bounded slice of the same workload.

    python bench.py [--gpus 1] [--steps 64] [--warmup 4] [--workload chart]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mpaths/s + p50/p99 first-frame latency, served vs plain, at relMSE 1e-3"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per unit of work, in the reference's record layout
# (crates/cenote/src/wavefront.rs:211-356, shaders/pathstate.slang): a build
# that compresses state is credited in time, not bytes.
BYTES = {
    "camera_path": 120,   # raygen 56 + zero-fill 16 + accumulate 48
    "camera_path_aov": 144,
    "ray": 60,            # trace (intersect part): state read 40 + hit/queue write 20
    "hit": 200 + 396,     # shade_surface: state 76 + write-back 56 + shadow 68; scene records 396
    "miss": 72 + 68,      # shade_miss: state 40 + radiance RMW 32; env texels + pdf 68
    "shadow": 96,         # trace (shadow part): record 64 + radiance RMW 32
}


def log(*args):
    print(*args, file=sys.stderr, flush=True)


def algorithmic_bytes(counts, traversal, scene_info, pixels, waves, aov):
    """Per-kernel algorithmic bytes over `waves` waves."""
    node, tri = scene_info.bvh_node_bytes, scene_info.bvh_triangle_bytes
    rays = int(counts["rays"].sum())
    hits = int(counts["hits"].sum())
    misses = int(counts["misses"].sum())
    shadows = int(counts["shadows"].sum())
    return {
        "raygen": pixels * waves * 72,
        "accumulate": pixels * waves * (BYTES["camera_path_aov"] if aov else 48),
        # intersect fused with the previous bounce's trace_shadow
        "trace": BYTES["ray"] * rays + node * traversal["ray_nodes"] + tri * traversal["ray_triangles"]
        + BYTES["shadow"] * shadows + node * traversal["shadow_nodes"] + tri * traversal["shadow_triangles"],
        "shade_surface": BYTES["hit"] * hits,
        "shade_miss": BYTES["miss"] * misses,
    }, dict(rays=rays, hits=hits, misses=misses, shadows=shadows)


def cpu_baseline(data, width, height, max_bounces, budget_s=12.0):
    """The CPU oracle (oracle/, a C++ restatement of the reference estimator)
    on the host's cores, over a bounded band of rows of the same workload."""
    from oracle import oracle as orc

    threads = min(16, orc.threads_default())
    scene = orc.OracleScene(data)
    # Probe a band of rows; then trace whole frames at successive sample
    # indices until budget_s has passed (or one band, if a frame would not fit).
    t0 = time.perf_counter()
    scene.trace_rows(width, height, 0, 16, 0, max_bounces=max_bounces, threads=threads)
    frame_s = (time.perf_counter() - t0) * height / 16
    if frame_s > budget_s:
        rows = max(16, int(height * budget_s / frame_s))
        row0 = (height - rows) // 2
        t0 = time.perf_counter()
        scene.trace_rows(width, height, row0, row0 + rows, 0, max_bounces=max_bounces, threads=threads)
        elapsed = time.perf_counter() - t0
        paths = rows * width
        sample = f"rows {row0}..{row0 + rows} of one {width}x{height} wave (sample 0)"
    else:
        waves = 0
        t0 = time.perf_counter()
        while waves == 0 or time.perf_counter() - t0 < budget_s:
            scene.trace(width, height, waves, max_bounces=max_bounces, threads=threads)
            waves += 1
        elapsed = time.perf_counter() - t0
        paths = waves * width * height
        sample = f"{waves} full {width}x{height} waves (samples 0..{waves - 1})"
    return {
        "value": round(paths / elapsed / 1e6, 4),
        "unit": "Mpaths/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample}, depth {max_bounces}: {paths} camera paths in {elapsed:.2f} s",
    }


def load_traffic(kernel: str, workload: str):
    path = ROOT / "profiles" / "traffic.json"
    if not path.exists():
        return None
    try:
        table = json.loads(path.read_text())
        entry = table.get(workload, {}).get(kernel)
        return None if entry is None else entry.get("bytes_per_launch")
    except (ValueError, AttributeError):
        return None


def copy_calibration(gib: float = 1.0, copies: int = 8):
    """A plain device-to-device copy kernel of known size, run under the same
    PMC counters as the hot path: the ratio of FETCH_SIZE / WRITE_SIZE to the
    copy's known bytes calibrates the counters' units."""
    import torch

    n = int(gib * (1 << 30)) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(copies):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "copies": copies}), flush=True)


def _scenes(ctx, names):
    import cenote_amd as cn
    from cenote_amd import workloads

    data = {n: workloads.build(n) for n in names}
    return data, {n: cn.Scene(ctx, d) for n, d in data.items()}


def _common(metric_value, unit, steps, warmup, seconds, config):
    return {
        "metric": METRIC, "value": metric_value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(seconds / max(1, steps) * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": config,
    }


def batched_bench(args):
    """64 seeded views of the chart and heavy scenes (128^2..1024^2, aspect
    1..16/9, 1..64 spp): the plain form — view after view, one sample per
    wave from the reference's 2^20-path pool, a synchronous wave each (the
    reference's Renderer::accumulate loop) — against one many-view batch
    (shared waves from the 2^27-path pool, launch shape from the
    autotuner). Both forms render bitwise the same films."""
    import cenote_amd as cn
    from cenote_amd import autotune, workloads

    ctx = cn.Context(0)
    data, scenes = _scenes(ctx, ("chart", "heavy"))
    views = workloads.batched_views({n: d.camera for n, d in data.items()})
    total_paths = sum(w * h * spp for _, w, h, spp, _, _ in views)
    log(f"batched: {len(views)} views, {total_paths / 1e6:.1f} M camera paths")

    # Plain: one view at a time, one sample per synchronous wave.
    plain_r = {d: cn.Renderer(ctx, capacity=cn.DEFAULT_CAPACITY, max_bounces=d) for d in (12, 16)}
    films = [cn.Film(ctx, w, h) for _, w, h, _, _, _ in views]
    warm = cn.Film(ctx, 64, 64)
    for name, depth in (("chart", 12), ("heavy", 16)):  # first launches, outside the timing
        plain_r[depth].accumulate(scenes[name], warm)
    t0 = time.perf_counter()
    for (name, w, h, spp, depth, cam), film in zip(views, films):
        film.reset()
        scenes[name].set_camera(cam)
        for _ in range(spp):
            plain_r[depth].accumulate(scenes[name], film)
    plain_s = time.perf_counter() - t0
    plain_sums = [f.sums()[0] for f in films]
    for r in plain_r.values():
        r.close()
    log(f"plain: {plain_s:.3f} s, {total_paths / plain_s / 1e6:.1f} Mpaths/s")

    # Batched: autotune on a representative view, then one render_batch.
    r = cn.Renderer(ctx, max_bounces=16)
    tuned = autotune.tune(r, scenes["heavy"], 1024, 576, samples=4, repeats=1)
    log(f"autotuned shape: {tuned['shape']}")
    r.reserve(r.capacity, aov=True)
    warm.reset()
    r.render_batch([(scenes["chart"], warm, data["chart"].camera, 1, 12)])
    for f in films:
        f.reset()
    r.synchronize()
    t0 = time.perf_counter()
    r.render_batch([(scenes[n], f, cam, spp, d) for (n, w, h, spp, d, cam), f in zip(views, films)])
    batched_s = time.perf_counter() - t0
    same = all(np.array_equal(f.sums()[0], p, equal_nan=True) for f, p in zip(films, plain_sums))
    log(f"batched: {batched_s:.3f} s, {total_paths / batched_s / 1e6:.1f} Mpaths/s; films identical: {same}")
    for f in films + [warm]:
        f.close()
    r.close()
    value = total_paths / batched_s / 1e6
    out = _common(round(value, 3), "Mpaths/s", 1, 1, batched_s, {
        "workload": f"batched: {len(views)} views of chart (depth 12) and heavy (depth 16), 128^2..1024^2, 1..64 spp",
        "paths": total_paths, "form": "one many-view batch vs the plain one-sample-per-wave loop",
        "shape": tuned["shape"]})
    out["plain"] = {"value": round(total_paths / plain_s / 1e6, 3), "unit": "Mpaths/s", "seconds": round(plain_s, 3)}
    out["speedup_vs_plain"] = round(plain_s / batched_s, 3)
    out["films_bitwise_equal"] = bool(same)
    return out


def serving_bench(args):
    """An open-loop Poisson stream of progressive requests over the resident
    Cornell, chart and heavy scenes (256^2 / 720p / 1080p; half target 1..256
    spp, half relMSE 1e-2 capped at 256 spp), served by the library's request
    loop, then replayed through the plain form with the samples each request
    needed — the same images, so the comparison is at matched noise."""
    import cenote_amd as cn
    from cenote_amd import serve, workloads

    ctx = cn.Context(0)
    data, scenes = _scenes(ctx, ("cornell", "chart", "heavy"))
    depths = {n: workloads.CONFIGS[n]["max_bounces"] for n in data}
    specs = serve.poisson_requests({n: d.camera for n, d in data.items()}, depths, args.rate, args.requests)
    log(f"serving: {len(specs)} requests at {args.rate} req/s over {specs[-1].arrival:.2f} s")
    r = cn.Renderer(ctx, max_bounces=16)
    r.reserve(1 << 24, aov=True)  # the server's waves hold at most 2^24 paths
    budget = 8 << 30
    sv = serve.Server(r, budget)
    # Warm: one small request through every scene.
    for n in scenes:
        sv.submit(scenes[n], data[n].camera, 64, 64, 1, max_bounces=depths[n])
    sv.drain()
    sv.close()
    sv = serve.Server(r, budget)
    served = serve.run_served(sv, scenes, specs)
    sv.close()
    r.close()
    s_sum = serve.summarize(served)
    log(f"served: {s_sum}")
    samples = [rec["samples"] for rec in served]
    plain = serve.run_plain(ctx, scenes, specs, samples)
    p_sum = serve.summarize(plain)
    log(f"plain: {p_sum}")
    out = _common(s_sum["mpaths_per_s"], "Mpaths/s", len(specs), 0, s_sum["seconds"], {
        "workload": f"serving: {len(specs)} Poisson requests at {args.rate}/s over cornell/chart/heavy, "
                    "256^2/720p/1080p, 1..256 spp or relMSE 1e-2",
        "hbm_budget_bytes": budget, "form": "continuous batching at sample granularity vs plain FIFO"})
    out["served"] = s_sum
    out["plain"] = p_sum
    out["speedup_vs_plain"] = round(s_sum["mpaths_per_s"] / p_sum["mpaths_per_s"], 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--workload", default="chart", choices=["cornell", "chart", "heavy", "batched", "serving"])
    ap.add_argument("--rate", type=float, default=10.0, help="serving: Poisson arrival rate, requests/s")
    ap.add_argument("--requests", type=int, default=48, help="serving: requests in the stream")
    ap.add_argument("--serving-requests", type=int, default=32,
                    help="chart: also serve this many Poisson requests, served vs plain (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--capacity", type=int, default=0, help="path-pool capacity (0: the library's default)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device copy (the PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the path shards by tile or "
              "view across GPUs, which this build does not run)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched":
        print(json.dumps(batched_bench(args)), flush=True)
        return
    if args.workload == "serving":
        print(json.dumps(serving_bench(args)), flush=True)
        return

    import cenote_amd as cn
    from cenote_amd import workloads

    cfg = workloads.CONFIGS[args.workload]
    width, height, depth, aov = cfg["width"], cfg["height"], cfg["max_bounces"], cfg["aov"]
    t0 = time.perf_counter()
    data = workloads.build(args.workload)
    gen_s = time.perf_counter() - t0
    log(f"workload {args.workload}: {data.triangle_count} triangles, generated in {gen_s:.1f} s")

    ctx = cn.Context(0)
    name, cus = ctx.describe()
    scene = cn.Scene(ctx, data)
    info = scene.info()
    log(f"device {name} ({cus} CUs); scene prep {info.prep_seconds:.2f} s (BVH {info.bvh_build_seconds:.2f} s, "
        f"{info.bvh_node_count} nodes, depth {info.bvh_max_depth})")
    renderer = cn.Renderer(ctx, capacity=args.capacity, max_bounces=depth)
    film = cn.Film(ctx, width, height)
    pixels = width * height
    steps = args.steps
    # Size the pool for the timed waves up front, so no wave allocates.
    renderer.reserve(min(steps, renderer.samples_per_wave(width, height)) * pixels, aov=True)

    # Counting pass: the same waves the timed region traces, with traversal
    # counters on, for the algorithmic-bytes model.
    renderer.set_counters(True)
    counts_total = None
    for s in range(steps):
        renderer.accumulate(scene, film)
        c = renderer.wave_counts()
        counts_total = c if counts_total is None else {k: counts_total[k] + c[k] for k in c}
    traversal = renderer.traversal_counts()
    renderer.set_counters(False)

    # Warmup, then the timed region: the same sample indices 0..steps-1.
    film.reset()
    renderer.accumulate(scene, film, count=args.warmup)
    film.reset()
    renderer.set_profiling(True)
    renderer.synchronize()
    t0 = time.perf_counter()
    renderer.accumulate(scene, film, count=steps, synchronize=False)
    renderer.synchronize()
    elapsed = time.perf_counter() - t0
    times = renderer.kernel_times()
    renderer.set_profiling(False)

    value = pixels * steps / elapsed / 1e6
    per_kernel_bytes, totals = algorithmic_bytes(counts_total, traversal, info, pixels, steps, aov)
    dominant = max(times, key=lambda k: times[k][0])
    dom_ms, dom_launches = times[dominant]
    dom_bytes = per_kernel_bytes.get(dominant, 0)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    grays = (totals["rays"] + totals["shadows"]) / elapsed / 1e9
    kernel_ms = {k: round(v[0] / steps, 4) for k, v in times.items()}
    kernel_gbs = {k: round(per_kernel_bytes[k] / (times[k][0] * 1e-3) / 1e9, 1)
                  for k in times if times[k][0] > 0 and k in per_kernel_bytes}
    log(f"per wave: rays {totals['rays'] / steps:.0f}, hits {totals['hits'] / steps:.0f}, "
        f"misses {totals['misses'] / steps:.0f}, shadows {totals['shadows'] / steps:.0f}; "
        f"nodes/ray {traversal['ray_nodes'] / max(1, totals['rays']):.1f}, "
        f"tris/ray {traversal['ray_triangles'] / max(1, totals['rays']):.1f}, "
        f"nodes/shadow {traversal['shadow_nodes'] / max(1, totals['shadows']):.1f}, "
        f"tris/shadow {traversal['shadow_triangles'] / max(1, totals['shadows']):.1f}")
    log(f"timed {steps} waves in {elapsed:.3f} s: {value:.1f} Mpaths/s, {grays:.3f} Grays/s")
    log(f"per-wave kernel ms: {kernel_ms}")
    log(f"per-kernel algorithmic GB/s: {kernel_gbs}")

    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload} {width}x{height}, {steps} spp timed, depth {depth}"
                        + (", thin lens 0.02" if args.workload == "chart" else "")
                        + (", AOVs" if aov else ""),
            "triangles": data.triangle_count,
            "pixels": pixels,
            "path_pool": renderer.capacity,
            "form": "plain (one progressive frame, one sample per step; waves of several samples)",
        },
        "grays_per_s": round(grays, 4),
        "kernel_ms_per_step": kernel_ms,
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(dominant, args.workload),
            "algorithmic_bytes_per_launch": int(dom_bytes / max(1, dom_launches)),
            "avg_launch_ms": round(dom_ms / max(1, dom_launches), 5),
        },
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(data, width, height, depth)
        log(f"cpu baseline: {result['cpu_baseline']}")
    film.close()
    renderer.close()
    scene.close()
    ctx.close()
    if args.workload == "chart" and args.serving_requests > 0:
        # The metric's other half: the same GPU serving a Poisson request
        # stream, continuous batching vs the plain one-frame-at-a-time loop.
        served = serving_bench(argparse.Namespace(rate=args.rate, requests=args.serving_requests))
        result["served_vs_plain"] = {k: served[k] for k in ("config", "served", "plain", "speedup_vs_plain")}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
