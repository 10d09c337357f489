"""Benchmark of the wavefront path-tracing hot path on one MI355X.

A step is one full frame of the workload: every pixel at the workload's
samples per pixel, traced to its depth cap and folded into a reset film
(Renderer::accumulate x spp, render/mod.rs:245-259). The default workload
is the 2M-triangle scene — 1920x1080, 32 spp, depth 16, beauty plus
albedo/normal/depth AOVs — and the line reports its camera paths per second
over K timed frames, the dominant kernel's algorithmic bandwidth against the
HBM roofline (HIP events on the launch stream) with the measured HBM
traffic of the same run shape beside it, and the CPU oracle timed on a
bounded slice of the same workload. Beside the headline, the same run
measures the material chart (1280x720, 64 spp), the 2M-triangle scene's
diagonal view (72% of camera rays hit geometry), the Cornell box, the
64-view batch against two plain forms, and the serving loop over a Poisson
request stream at 10, 50, 100 and 200 requests/s against two plain forms.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--workload heavy]
    python bench.py --workload batched | serving        (one leg alone)
    python bench.py --reduce DIR --workload heavy --steps 2   (rocprofv3 csv -> per-launch table)
    python bench.py --reduce-traffic DIR --workload heavy --steps 3   (PMC runs -> measured bytes per launch)
"""

from __future__ import annotations

import argparse
import csv
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mpaths/s + p50/p99 first-frame latency, served vs plain, at relMSE 1e-3"
HBM_PEAK_GBS = 8000.0
KERNELS = ("raygen", "trace", "shade_miss", "shade_surface", "accumulate", "tail")
KERNEL_SYMBOLS = {"raygen": "k_raygen", "trace": "k_trace", "shade_miss": "k_shade_miss",
                  "shade_surface": "k_shade_surface", "accumulate": "k_accumulate", "tail": "k_tail"}

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


# --- host ----------------------------------------------------------------------


def host_cpus() -> dict:
    """The host's CPUs as this process sees them: os.cpu_count(), the
    affinity mask, and the cgroup CPU quota (cpu.max) when one is set. The
    oracle runs one thread per CPU of the smaller of the last two."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    quota = None
    for path in ("/sys/fs/cgroup/cpu.max",):
        try:
            text = Path(path).read_text().split()
            if text and text[0] != "max":
                quota = float(text[0]) / float(text[1])
        except (OSError, ValueError, IndexError):
            pass
    threads = affinity if quota is None else max(1, min(affinity, int(math.ceil(quota))))
    return {"nproc": nproc, "affinity": affinity, "cgroup_cpu_quota": quota, "threads": threads}


def cpu_baseline(data, width, height, spp, max_bounces, budget_s=12.0):
    """The CPU oracle (oracle/, a C++ restatement of the reference estimator)
    on every CPU this process may use, over a bounded band of rows of the
    same workload at its own sample indices."""
    from oracle import oracle as orc

    host = host_cpus()
    threads = host["threads"]
    scene = orc.OracleScene(data)
    t0 = time.perf_counter()
    scene.trace_rows(width, height, 0, 8, 0, max_bounces=max_bounces, threads=threads)
    band_s = time.perf_counter() - t0
    frame_sample_s = band_s * height / 8
    if frame_sample_s > budget_s:
        rows = max(8, min(height, int(height * budget_s / frame_sample_s)))
        row0 = (height - rows) // 2
        t0 = time.perf_counter()
        scene.trace_rows(width, height, row0, row0 + rows, 0, max_bounces=max_bounces, threads=threads)
        elapsed = time.perf_counter() - t0
        paths = rows * width
        sample = f"rows {row0}..{row0 + rows} of a {width}x{height} frame at sample 0"
    else:
        samples = 0
        t0 = time.perf_counter()
        while samples < spp and (samples == 0 or time.perf_counter() - t0 < budget_s):
            scene.trace(width, height, samples, max_bounces=max_bounces, threads=threads)
            samples += 1
        elapsed = time.perf_counter() - t0
        paths = samples * width * height
        sample = f"{samples} of the frame's {spp} samples of every {width}x{height} pixel"
    return {
        "value": round(paths / elapsed / 1e6, 4),
        "unit": "Mpaths/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample}, depth {max_bounces}: {paths} camera paths in {elapsed:.2f} s",
        "host": host,
    }


# --- the plain path ------------------------------------------------------------


def waves_per_frame(width, height, spp, capacity):
    per = max(1, capacity // (width * height))
    return -(-spp // per)


def tail_split(depth, tail):
    """The bounce the tail launch takes over after (its trace is the last
    wavefront trace), or `depth` when every bounce runs as stages."""
    return tail if 0 < tail < depth else depth


def timed_launches(name, steps, capacity, tail=0):
    """Dispatches of each stage kernel in the timed region of `steps` frames."""
    from cenote_amd import workloads

    cfg = workloads.CONFIGS[name]
    waves = waves_per_frame(cfg["width"], cfg["height"], cfg["spp"], capacity) * steps
    depth = cfg["max_bounces"]
    t = tail_split(depth, tail)
    traced = waves * min(depth, t + 1)
    # bounce 0's misses are shaded by its trace launch
    return {"raygen": waves, "trace": traced, "shade_miss": waves * max(0, t - 1), "shade_surface": waves * t,
            "accumulate": waves, "tail": waves if t < depth else 0}


def shape_key(name, capacity, width=4, tail=0):
    """The run shape a measured-traffic entry belongs to."""
    from cenote_amd import workloads

    cfg = workloads.CONFIGS[name]
    spw = min(cfg["spp"], max(1, capacity // (cfg["width"] * cfg["height"])))
    t = tail_split(cfg["max_bounces"], tail)
    return (f"{name} {cfg['width']}x{cfg['height']} spp {cfg['spp']} depth {cfg['max_bounces']} "
            f"pool {capacity} samples/wave {spw} bvh {width}-wide"
            + (f" tail after bounce {t}" if t < cfg["max_bounces"] else ""))


def load_traffic(key):
    path = ROOT / "profiles" / "traffic.json"
    try:
        table = json.loads(path.read_text())
    except (OSError, ValueError):
        return None
    entry = table.get("runs", {}).get(key)
    return entry if isinstance(entry, dict) else None


def algorithmic_bytes(counts, by_bounce, info, pixels, aov, width=4, tail=0, counter_lines=0):
    """Per-kernel algorithmic bytes of the counted waves, bounce by bounce; a
    node fetch counts its record size (64 bytes 4-wide, 128 bytes 8-wide).
    With a tail after bounce t, trace holds bounces 0..t (and the camera
    paths' generation and bounce 0's miss shading, which its bounce-0 launch
    does: the reference's raygen stage, 72 bytes per path), shade_miss
    bounces 1..t-1, shade_surface bounces 0..t-1, and the tail kernel every
    later trace and all shading from bounce t on. raygen itself only resets
    the range's queue counters (128 bytes a counter line). `by_bounce[b]` = ray nodes, ray triangles,
    shadow nodes, shadow triangles of bounce b's trace launch (whose shadow
    rays are bounce b-1's connections)."""
    node = 128 if width == 8 else info.bvh_node_bytes
    tri = info.bvh_triangle_bytes
    depth = len(counts["rays"])
    traced, surface, miss = [], [], []
    for b in range(depth):
        rn, rt, sn, st = (int(v) for v in by_bounce[b])
        connections = int(counts["shadows"][b - 1]) if b > 0 else 0
        # intersect fused with the previous bounce's trace_shadow
        traced.append(BYTES["ray"] * int(counts["rays"][b]) + node * rn + tri * rt
                      + BYTES["shadow"] * connections + node * sn + tri * st)
        surface.append(BYTES["hit"] * int(counts["hits"][b]))
        miss.append(BYTES["miss"] * int(counts["misses"][b]))
    t = tail_split(depth, tail)
    out = {
        "raygen": counter_lines * 128,
        "accumulate": pixels * (BYTES["camera_path_aov"] if aov else 48),
        "trace": sum(traced[: t + 1]) + miss[0] + pixels * 72,
        "shade_surface": sum(surface[:t]),
        "shade_miss": sum(miss[1:t]),
    }
    if t < depth:
        out["tail"] = sum(traced[t + 1:]) + sum(surface[t:]) + sum(miss[t:])
    totals = {k: int(counts[k].sum()) for k in ("rays", "hits", "misses", "shadows")}
    return out, totals


def measure_plain(ctx, name, steps, warmup, capacity=0, data=None, cpu=False, bvh_width=0, tail_bounce=None,
                  shape=None):
    """`steps` full frames of workload `name`, each a reset film and the
    workload's spp, stream-ordered with one synchronize at the end."""
    import cenote_amd as cn
    from cenote_amd import workloads

    cfg = workloads.CONFIGS[name]
    width, height, spp, depth, aov = cfg["width"], cfg["height"], cfg["spp"], cfg["max_bounces"], cfg["aov"]
    t0 = time.perf_counter()
    if data is None:
        data = workloads.build(name)
    gen_s = time.perf_counter() - t0
    scene = cn.Scene(ctx, data)
    info = scene.info()
    log(f"{name}: {data.triangle_count} triangles (generated in {gen_s:.1f} s); scene prep {info.prep_seconds:.2f} s "
        f"(BVH {info.bvh_build_seconds:.2f} s, {info.bvh_node_count} nodes, depth {info.bvh_max_depth})")
    renderer = cn.Renderer(ctx, capacity=capacity, max_bounces=depth)
    if bvh_width:
        renderer.shape = {"bvh_width": bvh_width}
    if tail_bounce is not None:
        renderer.shape = {"tail_bounce": tail_bounce}
    if shape:
        renderer.shape = shape
    # the tree the traversal walks (a scene without an 8-wide tree is walked 4-wide)
    walked = 8 if renderer.shape["bvh_width"] == 8 and info.bvh8_node_count else 4
    film = cn.Film(ctx, width, height)
    pixels = width * height
    per_wave = min(spp, renderer.samples_per_wave(width, height))
    renderer.reserve(per_wave * pixels, aov=True)

    # Counting pass: one frame (every frame traces the same samples), with
    # traversal counters on, for the algorithmic-bytes model.
    renderer.set_counters(True)
    counts = None
    film.reset()
    done = 0
    while done < spp:
        n = min(per_wave, spp - done)
        renderer.accumulate(scene, film, count=n)
        c = renderer.wave_counts()
        counts = c if counts is None else {k: counts[k] + c[k] for k in c}
        done += n
    traversal = renderer.traversal_counts()
    by_bounce = renderer.traversal_counts_by_bounce()
    renderer.set_counters(False)
    tail = renderer.shape["tail_bounce"]

    def frames(k):
        for _ in range(k):
            film.reset()
            renderer.accumulate(scene, film, count=spp, synchronize=False)

    frames(warmup)
    renderer.synchronize()
    renderer.set_profiling(True)
    t0 = time.perf_counter()
    frames(steps)
    renderer.synchronize()
    elapsed = time.perf_counter() - t0
    times = renderer.kernel_times()
    renderer.set_profiling(False)

    # the ranges' counter lines: 4 queues x 8 sub-queues per bounce, plus a spare
    # bounce, then the tail launch's 8 item counters
    lines = ((depth + 1) * 4 * 8 + 8) * waves_per_frame(width, height, spp, renderer.capacity)
    frame_bytes, totals = algorithmic_bytes(counts, by_bounce, info, pixels * spp, aov, walked, tail, lines)
    value = pixels * spp * steps / elapsed / 1e6
    grays = (totals["rays"] + totals["shadows"]) * steps / elapsed / 1e9
    kernel_ms = {k: round(v[0] / steps, 4) for k, v in times.items()}
    kernel_gbs = {k: round(frame_bytes[k] * steps / (times[k][0] * 1e-3) / 1e9, 1)
                  for k in times if times[k][0] > 0 and k in frame_bytes}
    dominant = max(times, key=lambda k: times[k][0])
    dom_ms, dom_launches = times[dominant]
    achieved = frame_bytes[dominant] * steps / (dom_ms * 1e-3) / 1e9
    avg_ms = dom_ms / max(1, dom_launches)
    key = shape_key(name, renderer.capacity, walked, tail)
    measured = load_traffic(key)
    traffic = None if measured is None else measured.get(dominant)
    hit_share = totals["hits"] and float(counts["hits"][0]) / float(counts["hits"][0] + counts["misses"][0])
    log(f"{name}: per frame rays {totals['rays']}, hits {totals['hits']}, misses {totals['misses']}, "
        f"shadows {totals['shadows']}; camera-ray hit share {hit_share:.3f}; "
        f"nodes/ray {traversal['ray_nodes'] / max(1, totals['rays']):.1f}, "
        f"tris/ray {traversal['ray_triangles'] / max(1, totals['rays']):.1f}, "
        f"nodes/shadow {traversal['shadow_nodes'] / max(1, totals['shadows']):.1f}")
    log(f"{name}: {steps} frames in {elapsed:.3f} s: {value:.1f} Mpaths/s, {grays:.3f} Grays/s; "
        f"kernel ms/frame {kernel_ms}; algorithmic GB/s {kernel_gbs}")
    roofline = {
        "bound": "hbm",
        "kernel": dominant,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None if traffic is None else int(traffic["bytes_per_launch"]),
        "basis": "achieved = algorithmic bytes (reference record layout, issued, no cache credit) per launch / "
                 "average launch time; traffic = rocprofv3-measured HBM bytes per launch of this run shape",
        "algorithmic_bytes_per_launch": int(frame_bytes[dominant] * steps / max(1, dom_launches)),
        "avg_launch_ms": round(avg_ms, 5),
        "launches": dom_launches,
    }
    if traffic is not None:
        measured_gbs = traffic["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9
        ratio = roofline["algorithmic_bytes_per_launch"] / traffic["bytes_per_launch"]
        roofline["traffic_gbs"] = round(measured_gbs, 1)
        roofline["traffic_frac"] = round(measured_gbs / HBM_PEAK_GBS, 4)
        roofline["traffic_profile"] = measured.get("_source")
        # Issued bytes over HBM bytes: above 3 the caches (L2, the 256 MB
        # Infinity Cache) absorb most of the issued traffic, and the kernel
        # is bound by issue and latency rather than by HBM.
        roofline["algorithmic_over_traffic"] = round(ratio, 2)
        roofline["cache_absorbed"] = ratio > 3.0
    result = {
        "workload": name,
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "ms_per_frame": round(elapsed / steps * 1e3, 4),
        "frames": steps,
        "seconds": round(elapsed, 4),
        "grays_per_s": round(grays, 4),
        "camera_hit_share": round(hit_share, 4),
        "kernel_ms_per_frame": kernel_ms,
        "roofline": roofline,
        "shape_key": key,
        "config": {
            "workload": f"{name} {width}x{height}, {spp} spp, depth {depth}"
                        + (", thin lens 0.02" if name == "chart" else "") + (", AOVs" if aov else ""),
            "triangles": data.triangle_count, "pixels": pixels, "spp": spp, "path_pool": renderer.capacity,
            "samples_per_wave": per_wave,
            "bvh_width": walked,
            "tail_after_bounce": tail_split(depth, tail) if tail_split(depth, tail) < depth else None,
        },
    }
    if cpu:
        result["cpu_baseline"] = cpu_baseline(data, width, height, spp, depth)
        log(f"{name}: cpu baseline {result['cpu_baseline']}")
    film.close()
    renderer.close()
    scene.close()
    return result


# --- batched and served forms ------------------------------------------------


def _scenes(ctx, names, data=None):
    import cenote_amd as cn
    from cenote_amd import workloads

    data = dict(data or {})
    for n in names:
        if n not in data:
            data[n] = workloads.build(n)
    return data, {n: cn.Scene(ctx, data[n]) for n in names}


def batched_bench(ctx, data=None):
    """64 seeded views of the chart and heavy scenes (128^2..1024^2, aspect
    1..16/9, 1..64 spp) three ways: the reference-shaped plain loop (view
    after view, one synchronous sample per wave from a 2^20-path pool, as
    Renderer::accumulate is called), the best plain loop (view after view,
    all of a view's samples stream-ordered from the 2^27-path pool), and one
    many-view batch (shared waves, launch shape from the autotuner). All
    three render bitwise the same films."""
    import cenote_amd as cn
    from cenote_amd import autotune, workloads

    data, scenes = _scenes(ctx, ("chart", "heavy"), data)
    views = workloads.batched_views({n: data[n].camera for n in scenes})
    total_paths = sum(w * h * spp for _, w, h, spp, _, _ in views)
    log(f"batched: {len(views)} views, {total_paths / 1e6:.1f} M camera paths")
    films = [cn.Film(ctx, w, h) for _, w, h, _, _, _ in views]
    warm = cn.Film(ctx, 64, 64)

    def plain(capacity, per_sample):
        rs = {d: cn.Renderer(ctx, capacity=capacity, max_bounces=d) for d in (12, 16)}
        for name, depth in (("chart", 12), ("heavy", 16)):  # first launches, outside the timing
            warm.reset()
            rs[depth].accumulate(scenes[name], warm)
        if not per_sample:
            biggest = max(w * h * spp for _, w, h, spp, _, _ in views)
            for r in rs.values():
                r.reserve(min(biggest, r.capacity), aov=True)
        t0 = time.perf_counter()
        for (name, w, h, spp, depth, cam), film in zip(views, films):
            film.reset()
            scenes[name].set_camera(cam)
            if per_sample:
                for _ in range(spp):
                    rs[depth].accumulate(scenes[name], film)
            else:
                rs[depth].accumulate(scenes[name], film, count=spp)
        seconds = time.perf_counter() - t0
        sums = [f.sums()[0] for f in films]
        for r in rs.values():
            r.close()
        return seconds, sums

    ref_s, ref_sums = plain(cn.DEFAULT_CAPACITY, True)
    log(f"batched: reference-shaped plain {ref_s:.3f} s, {total_paths / ref_s / 1e6:.1f} Mpaths/s")
    best_s, best_sums = plain(0, False)
    log(f"batched: best plain {best_s:.3f} s, {total_paths / best_s / 1e6:.1f} Mpaths/s")

    r = cn.Renderer(ctx, max_bounces=16)
    tuned = autotune.tune(r, scenes["heavy"], 1024, 576, samples=4, repeats=1)
    log(f"batched: autotuned shape {tuned['shape']}")
    r.reserve(r.capacity, aov=True)
    warm.reset()
    r.render_batch([(scenes["chart"], warm, data["chart"].camera, 1, 12)])
    for f in films:
        f.reset()
    r.synchronize()
    t0 = time.perf_counter()
    r.render_batch([(scenes[n], f, cam, spp, d) for (n, w, h, spp, d, cam), f in zip(views, films)])
    batched_s = time.perf_counter() - t0
    sums = [f.sums()[0] for f in films]
    same = all(np.array_equal(a, b, equal_nan=True) and np.array_equal(a, c, equal_nan=True)
               for a, b, c in zip(sums, ref_sums, best_sums))
    log(f"batched: one batch {batched_s:.3f} s, {total_paths / batched_s / 1e6:.1f} Mpaths/s; films identical: {same}")
    for f in films + [warm]:
        f.close()
    r.close()
    for s in scenes.values():
        s.close()
    rate = lambda s: round(total_paths / s / 1e6, 3)  # noqa: E731
    return {
        "workload": f"batched: {len(views)} views of chart (depth 12) and heavy (depth 16), 128^2..1024^2, 1..64 spp",
        "paths": total_paths,
        "batched": {"value": rate(batched_s), "unit": "Mpaths/s", "seconds": round(batched_s, 4),
                    "shape": tuned["shape"]},
        "plain_reference_shaped": {"value": rate(ref_s), "unit": "Mpaths/s", "seconds": round(ref_s, 4),
                                   "form": "view by view, one synchronous sample per wave, 2^20-path pool"},
        "plain_best": {"value": rate(best_s), "unit": "Mpaths/s", "seconds": round(best_s, 4),
                       "form": "view by view, all of a view's samples stream-ordered, 2^27-path pool"},
        "speedup_vs_reference_shaped": round(ref_s / batched_s, 3),
        "speedup_vs_best_plain": round(best_s / batched_s, 3),
        "films_bitwise_equal": bool(same),
    }


def serving_bench(ctx, rates=(10.0, 50.0, 100.0, 200.0), requests=48, data=None):
    """An open-loop Poisson stream of progressive requests over the resident
    Cornell, chart and heavy scenes (256^2 / 720p / 1080p; half target
    1..256 spp, half relMSE 1e-2 capped at 256 spp), served by the library's
    request loop at each arrival rate, against two plain forms run on the
    same requests with the samples each request needed (so the images are
    the same): the reference-shaped loop (one synchronous sample per wave
    from a 2^20-path pool) and the best plain loop (the 2^27-path pool, the
    first sample as its own wave for the first frame, the rest
    stream-ordered). Each plain form runs its requests back to back once;
    its latencies under each arrival schedule follow from first-in
    first-out service with those measured service times."""
    import cenote_amd as cn
    from cenote_amd import serve, workloads

    data, scenes = _scenes(ctx, ("cornell", "chart", "heavy"), data)
    depths = {n: workloads.CONFIGS[n]["max_bounces"] for n in scenes}
    cams = {n: data[n].camera for n in scenes}
    r = cn.Renderer(ctx, max_bounces=16)  # depth and launch shape for the server's lanes
    budget = 8 << 30
    sv = serve.Server(r, budget)
    for n in scenes:  # warm: one small request through every scene
        sv.submit(scenes[n], data[n].camera, 64, 64, 1, max_bounces=depths[n])
    sv.drain()
    sv.close()
    served = {}
    samples = None
    for rate in rates:
        specs = serve.poisson_requests(cams, depths, rate, requests)
        sv = serve.Server(r, budget)
        records = serve.run_served(sv, scenes, specs)
        sv.close()
        summary = serve.summarize(records)
        summary["arrival_span_s"] = round(specs[-1].arrival, 4)
        served[rate] = (specs, records, summary)
        if samples is None:
            samples = [rec["samples"] for rec in records]
        log(f"serving: served at {rate:g} req/s: {summary}")
    r.close()
    specs0 = served[rates[0]][0]
    service = {
        "plain_reference_shaped": serve.plain_service(ctx, scenes, specs0, samples, capacity=cn.DEFAULT_CAPACITY,
                                                      per_sample=True),
        "plain_best": serve.plain_service(ctx, scenes, specs0, samples, capacity=0, per_sample=False),
    }
    out = {"workload": f"serving: {requests} Poisson requests over cornell/chart/heavy, 256^2/720p/1080p, "
                       "1..256 spp or relMSE 1e-2", "hbm_budget_bytes": budget,
           "form": "continuous batching at sample granularity vs one request at a time", "rates": {}}
    for rate in rates:
        specs, records, summary = served[rate]
        # The plain forms render each request with the samples the first
        # served run gave it; served runs at other rates may stop a noise
        # target a wave later or earlier, so each form is compared per path.
        row = {"served": summary}
        for form, svc in service.items():
            p = serve.summarize(serve.fifo_schedule(specs, svc))
            row[form] = p
            row[f"speedup_vs_{form[6:]}"] = round(summary["mpaths_per_s"] / p["mpaths_per_s"], 3)
        out["rates"][f"{rate:g}"] = row
        log(f"serving: {rate:g} req/s: " + json.dumps({k: v for k, v in row.items() if k.startswith("speedup")}))
    for form, svc in service.items():
        busy = sum(s["service"] for s in svc)
        paths = sum(s["paths"] for s in svc)
        out[f"{form}_saturated_mpaths_per_s"] = round(paths / busy / 1e6, 3)
    for s in scenes.values():
        s.close()
    return out


# --- profiles ----------------------------------------------------------------------


def copy_calibration(gib: float = 1.0, copies: int = 8):
    """A plain device-to-device copy of known size, run under the same PMC
    counters as the hot path: the ratio of FETCH_SIZE / WRITE_SIZE to the
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


CALIBRATION_KINDS = {"fma_f32": 0, "add_u32": 1, "max_f32": 2}


def valu_calibration(iterations: int = 4096, blocks_per_cu: int = 32):
    """Kernels of known VALU instruction count (cn_calibrate_valu: eight
    independent chains of f32 FMA, u32 add or f32 max per lane, every SIMD
    full), timed with events and run under the same SQ counters as the hot
    path: the fastest kind's counted VALU instructions per busy cycle is the
    issue rate the stage kernels' are read against. Dispatches come in pairs
    (warm, timed) in the order of CALIBRATION_KINDS."""
    import ctypes as C

    import cenote_amd as cn
    from cenote_amd import _abi

    ctx = cn.Context(0)
    _, cus = ctx.describe()
    blocks = cus * blocks_per_cu
    waves = blocks * 4
    out = {}
    for name, kind in CALIBRATION_KINDS.items():
        ms, per_wave = C.c_double(), C.c_uint64()
        _abi.check(ctx._lib.cn_calibrate_valu(ctx.handle, kind, iterations, blocks, C.byref(ms), C.byref(per_wave)))
        out[name] = {"iterations": iterations, "blocks": blocks, "waves": waves, "valu_per_wave": per_wave.value,
                     "ms": ms.value, "valu_per_s": waves * per_wave.value / (ms.value * 1e-3)}
    print(json.dumps({"valu_calibration": out}), flush=True)
    ctx.close()


def _csv_rows(directory, suffix):
    rows = []
    for path in sorted(Path(directory).rglob(f"*{suffix}")):
        with open(path, newline="") as f:
            rows.extend(csv.DictReader(f))
    return rows


def _base_name(kernel_name):
    """A demangled kernel name without its parameter list."""
    return kernel_name.replace("(anonymous namespace)::", "").split("(")[0]


def _stage_of(kernel_name):
    """The stage a dispatched kernel belongs to, or None. Counting-pass
    traversal kernels (COUNT = true) belong to no stage."""
    base = _base_name(kernel_name)
    for stage, sym in KERNEL_SYMBOLS.items():
        if sym + "<" in base or base.endswith(sym):
            if stage == "trace" and (sym + "<true") in base:
                return None
            return stage
    return None


def _short(kernel_name):
    return _base_name(kernel_name).replace("void ", "").replace("cn::", "").strip()


def reduce_profile(directory, name, steps, capacity, tail=0):
    """The timed region's dispatches of each stage kernel in a rocprofv3
    output directory of `bench.py --workload name --steps steps --warmup W
    --main-only`: per-launch average duration (kernel trace) or counter
    values (PMC), over the last dispatches of each stage kernel (the timed
    frames come last); the first timed frame's trace dispatches one by one
    (its bounces, in order); and every other kernel (scene build, the
    calibration kernels) averaged over all its dispatches."""
    launches = timed_launches(name, steps, capacity, tail)
    out, others, by_bounce = {}, {}, []
    per_dispatch = {}  # dispatch id -> {"kernel", "stage", "us", counters...}
    for r in _csv_rows(directory, "kernel_trace.csv"):
        d = per_dispatch.setdefault(int(r["Dispatch_Id"]), {"kernel": r["Kernel_Name"]})
        d["us"] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
    for r in _csv_rows(directory, "counter_collection.csv"):
        d = per_dispatch.setdefault(int(r["Dispatch_Id"]), {"kernel": r["Kernel_Name"]})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    stages = {k: [] for k in KERNEL_SYMBOLS}
    for did in sorted(per_dispatch):
        d = per_dispatch[did]
        stage = _stage_of(d["kernel"])
        if stage is not None:
            stages[stage].append(d)
        elif "k_trace<true" not in d["kernel"]:
            others.setdefault(_short(d["kernel"]), []).append(d)

    def averages(ds):
        keys = sorted({k for d in ds for k in d if k != "kernel"})
        return {("avg_us" if k == "us" else f"{k}_avg"): round(sum(d.get(k, 0.0) for d in ds) / len(ds), 3)
                for k in keys}

    for stage, ds in stages.items():
        timed = ds[-launches[stage]:] if 0 < launches[stage] <= len(ds) else []
        out[stage] = {"dispatches_all": len(ds), "dispatches_timed": len(timed), **(averages(timed) if timed else {})}
        if stage == "trace" and timed:
            per_frame = launches["trace"] // max(1, steps)
            by_bounce = [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items() if k != "kernel"}
                         for d in timed[:per_frame]]
    return {"kernels": out, "trace_first_frame": by_bounce,
            "others": {k: {"dispatches": len(v), **averages(v)} for k, v in others.items()}}


def _counter_per_launch(directory, counter):
    """Per-dispatch sums of one PMC counter in a rocprofv3 output directory,
    in dispatch order, with the kernel names."""
    per = {}
    for r in _csv_rows(directory, "counter_collection.csv"):
        if r["Counter_Name"] != counter:
            continue
        d = per.setdefault(int(r["Dispatch_Id"]), [r["Kernel_Name"], 0.0])
        d[1] += float(r["Counter_Value"])
    return [per[k] for k in sorted(per)]


def reduce_traffic(prof, name, steps, capacity, width=4, tail=0):
    """Measured HBM bytes per launch of each stage kernel, from the FETCH_SIZE
    and WRITE_SIZE runs of `bench.py --workload name --steps steps --main-only`
    (prof/pmc_FETCH_SIZE_<name>, prof/pmc_WRITE_SIZE_<name>), each counter
    scaled by what it reads for a 1 GiB device copy (`bench.py
    --copy-calibration` under the same counter: prof/pmc_*_copy): the copy's
    known bytes over the counter's KiB."""
    prof = Path(prof)
    factors = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        copies = [v for k, v in _counter_per_launch(prof / f"pmc_{counter}_copy", counter) if "copyBuffer" in k]
        factors[counter] = (1 << 30) / (sum(copies) / len(copies) * 1024.0)
    fetch = reduce_profile(prof / f"pmc_FETCH_SIZE_{name}", name, steps, capacity, tail)["kernels"]
    write = reduce_profile(prof / f"pmc_WRITE_SIZE_{name}", name, steps, capacity, tail)["kernels"]
    entry = {}
    for stage in KERNEL_SYMBOLS:
        f_kib, w_kib = fetch[stage].get("FETCH_SIZE_avg"), write[stage].get("WRITE_SIZE_avg")
        if f_kib is None or w_kib is None:
            continue
        entry[stage] = {"fetch_kib": f_kib, "write_kib": w_kib,
                        "bytes_per_launch": int(round((f_kib * factors["FETCH_SIZE"] + w_kib * factors["WRITE_SIZE"])
                                                      * 1024.0))}
    entry["_factors"] = {k: round(v, 4) for k, v in factors.items()}
    entry["_source"] = (f"rocprofv3 --pmc FETCH_SIZE | WRITE_SIZE -- python3 bench.py --workload {name} --steps {steps} "
                        f"--warmup 1 --main-only --no-cpu-baseline; calibrated by bench.py --copy-calibration")
    return {shape_key(name, capacity, width, tail): entry}


# --- main ----------------------------------------------------------------------


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed full frames")
    ap.add_argument("--warmup", type=int, default=2, help="untimed frames before the timed ones")
    ap.add_argument("--workload", default="heavy",
                    choices=["cornell", "chart", "heavy", "heavy_field", "batched", "serving"])
    ap.add_argument("--main-only", action="store_true", help="only the headline workload (profiling runs)")
    ap.add_argument("--rates", default="10,50,100,200", help="serving: Poisson arrival rates, requests/s")
    ap.add_argument("--requests", type=int, default=48, help="serving: requests in each stream")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--capacity", type=int, default=0, help="path-pool capacity (0: the library's default)")
    ap.add_argument("--bvh-width", type=int, default=0, choices=[0, 4, 8],
                    help="the tree the traversal walks (0: the library's default)")
    ap.add_argument("--tail-bounce", type=int, default=None,
                    help="the bounce after whose trace one launch finishes every live path (0: none; "
                         "default: the library's)")
    ap.add_argument("--shape", type=json.loads, default=None,
                    help='launch-shape fields to set, as JSON (e.g. \'{"shade_variant": 3}\')')
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device copy (the PMC byte-counter calibration)")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="run only the known-instruction-count kernel (the SQ VALU counter calibration)")
    ap.add_argument("--reduce", metavar="DIR", help="reduce a rocprofv3 output directory of this command's run")
    ap.add_argument("--reduce-traffic", metavar="DIR",
                    help="measured HBM bytes per launch from DIR/pmc_{FETCH,WRITE}_SIZE_{workload,copy}")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.valu_calibration:
        valu_calibration()
        return
    if args.reduce_traffic:
        from cenote_amd import _abi

        cap = args.capacity or _abi.POOL_CAPACITY
        tail = _abi.DEFAULT_TAIL_BOUNCE if args.tail_bounce is None else args.tail_bounce
        print(json.dumps(reduce_traffic(args.reduce_traffic, args.workload, args.steps, cap, args.bvh_width or 4, tail),
                         indent=1))
        return
    if args.reduce:
        from cenote_amd import _abi

        cap = args.capacity or _abi.POOL_CAPACITY
        tail = _abi.DEFAULT_TAIL_BOUNCE if args.tail_bounce is None else args.tail_bounce
        print(json.dumps({"shape_key": shape_key(args.workload, cap, args.bvh_width or 4, tail), "steps": args.steps,
                          **reduce_profile(args.reduce, args.workload, args.steps, cap, tail)}, indent=1))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the path shards by tile or "
              "view across GPUs, which this build does not run)", file=sys.stderr)
        sys.exit(2)

    import cenote_amd as cn

    ctx = cn.Context(0)
    name, cus = ctx.describe()
    log(f"device {name} ({cus} CUs)")
    rates = tuple(float(x) for x in args.rates.split(",") if x)

    if args.workload == "batched":
        out = batched_bench(ctx)
        print(json.dumps(out), flush=True)
        return
    if args.workload == "serving":
        out = serving_bench(ctx, rates, args.requests)
        print(json.dumps(out), flush=True)
        return

    from cenote_amd import workloads

    data = {args.workload: workloads.build(args.workload)}
    main_run = measure_plain(ctx, args.workload, args.steps, args.warmup, args.capacity, data[args.workload],
                             bvh_width=args.bvh_width, tail_bounce=args.tail_bounce, shape=args.shape,
                             cpu=not args.no_cpu_baseline)
    result = {
        "metric": METRIC,
        "value": main_run["value"],
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_run["ms_per_frame"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {**main_run["config"], "form": "plain: one full frame per step (reset film, every sample "
                                                  "stream-ordered), 2^27-path pool"},
        "grays_per_s": main_run["grays_per_s"],
        "camera_hit_share": main_run["camera_hit_share"],
        "kernel_ms_per_step": main_run["kernel_ms_per_frame"],
        "roofline": main_run["roofline"],
    }
    if "cpu_baseline" in main_run:
        result["cpu_baseline"] = main_run["cpu_baseline"]
    if not args.main_only:
        beside = {}
        for other in ("heavy_field", "chart", "cornell"):
            if other == args.workload:
                continue
            d = workloads.build(other)
            run = measure_plain(ctx, other, args.steps, args.warmup, args.capacity, d, bvh_width=args.bvh_width,
                                tail_bounce=args.tail_bounce, shape=args.shape)
            beside[other] = {k: run[k] for k in ("value", "unit", "ms_per_frame", "frames", "grays_per_s",
                                                 "camera_hit_share", "kernel_ms_per_frame", "roofline", "config")}
            if other in ("chart", "cornell"):
                data[other] = d
        result["plain_beside"] = beside
        shared = {k: v for k, v in data.items() if k in ("cornell", "chart", "heavy")}
        result["batched_vs_plain"] = batched_bench(ctx, shared)
        result["served_vs_plain"] = serving_bench(ctx, rates, args.requests, shared)
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
