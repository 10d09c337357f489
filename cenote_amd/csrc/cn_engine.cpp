// cn_engine.cpp — the C ABI's renderer and film halves: the wavefront
// engine (Wavefront::new / trace_then, wavefront.rs:486-828), the film
// (render/film.rs), and many-view batches.
//
// A wave is a list of jobs — (view, consecutive samples) pairs — laid end
// to end in a slot space. Jobs are grouped by scene; each group is walked
// in path-pool-sized ranges (raygen, then `max_bounces` rounds), all
// stream-ordered with no host round trip. One accumulate launch then folds
// every job's slots into its film in sample order, so the film's f32 sums
// are bit for bit those of one sample per wave (accumulate.slang:38-44).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "cn_handles.h"

using namespace cn;

namespace cn {

hipEvent_t EventProfiler::next() {
    if (used == pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        pool.push_back(e);
    }
    return pool[used++];
}

void EventProfiler::begin(int kind) {
    hipEvent_t e = next();
    if (!e) return;
    (void)hipEventRecord(e, stream);
    open.emplace_back(kind, used - 1);
}

void EventProfiler::end(int kind) {
    if (open.empty()) return;
    const auto [k, start] = open.back();
    open.pop_back();
    hipEvent_t e = next();
    if (!e) return;
    (void)hipEventRecord(e, stream);
    pairs.emplace_back(kind == k ? k : kind, start, used - 1);
}

void EventProfiler::collect() {
    for (const auto& [kind, a, b] : pairs) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, pool[a], pool[b]) == hipSuccess) {
            total_ms[kind] += ms;
            launches[kind] += 1;
        }
    }
    pairs.clear();
    open.clear();
    used = 0;
}

EventProfiler::~EventProfiler() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

}  // namespace cn

namespace cn {

// The path pool grows on demand up to the renderer's capacity, so a large
// capacity costs HBM only when views large enough to use it arrive.
int ensure_pool(cn_renderer* r, size_t paths) {
    paths = std::min<size_t>(paths, r->capacity);
    if (r->origin.count >= paths) return CN_OK;
    CN_HIP(hipStreamSynchronize(r->stream));
    r->drop_graphs();
    CN_HIP(r->origin.alloc(paths));
    CN_HIP(r->direction.alloc(paths));
    CN_HIP(r->throughput.alloc(paths));
    CN_HIP(r->hit.alloc(paths));
    CN_HIP(r->key.alloc(paths));
    CN_HIP(r->state.alloc(paths));
    // Every queue (and the shadow records) holds QUEUE_SPLIT sub-queues, each
    // sized for the most entries a pool-sized input can send it.
    const uint32_t sub = sub_capacity_for(uint32_t(paths));
    const size_t entries = size_t(sub) * QUEUE_SPLIT;
    CN_HIP(r->ray_queue0.alloc(entries));
    CN_HIP(r->ray_queue1.alloc(entries));
    CN_HIP(r->hit_queue.alloc(entries));
    CN_HIP(r->miss_queue.alloc(entries));
    CN_HIP(r->shadow_origin.alloc(entries));
    CN_HIP(r->shadow_dir.alloc(entries));
    CN_HIP(r->shadow_contrib.alloc(entries));
    CN_HIP(r->shadow_prim.alloc(entries));
    PathPool& p = r->pool;
    p.sub_capacity = sub;
    p.origin = r->origin.ptr;
    p.direction = r->direction.ptr;
    p.hit = r->hit.ptr;
    p.throughput = r->throughput.ptr;
    p.state = r->state.ptr;
    p.key = r->key.ptr;
    p.ray_queue[0] = r->ray_queue0.ptr;
    p.ray_queue[1] = r->ray_queue1.ptr;
    p.hit_queue = r->hit_queue.ptr;
    p.miss_queue = r->miss_queue.ptr;
    p.shadow_origin = r->shadow_origin.ptr;
    p.shadow_dir = r->shadow_dir.ptr;
    p.shadow_contrib = r->shadow_contrib.ptr;
    p.shadow_prim = r->shadow_prim.ptr;
    return CN_OK;
}

int ensure_slots(cn_renderer* r, size_t slots, bool aov) {
    const bool grow = r->radiance.count < slots || (aov && r->albedo.count < slots);
    if (!grow) return CN_OK;
    CN_HIP(hipStreamSynchronize(r->stream));  // nothing in flight may still use the old buffers
    r->drop_graphs();
    const size_t n = std::max(slots, r->radiance.count);
    if (r->radiance.count < n) CN_HIP(r->radiance.alloc(n));
    if (aov) {
        CN_HIP(r->albedo.alloc(n));
        CN_HIP(r->normal.alloc(n));
        CN_HIP(r->guide.alloc(n));
        CN_HIP(r->depth.alloc(n));
    }
    return CN_OK;
}

int enqueue_wave(cn_renderer* r, std::vector<HostJob> jobs, bool aov, const WaveArena* arena,
                 const hipEvent_t* after, uint32_t after_count) {
    if (jobs.empty()) return CN_OK;
    if (jobs.size() > MAX_JOBS) return fail(CN_E_INVALID, "a wave holds at most %u jobs", MAX_JOBS);
    for (HostJob& j : jobs) {
        if (j.max_bounces == 0) j.max_bounces = r->max_bounces;
        if (j.max_bounces > r->max_bounces)
            return fail(CN_E_INVALID, "a view's depth %u exceeds the renderer's %u", j.max_bounces, r->max_bounces);
    }
    // Group by scene, keeping first-appearance order (a scene's group walks
    // one contiguous slot span, so one SceneView serves each range).
    {
        std::vector<cn_scene*> order;
        std::vector<std::pair<size_t, size_t>> rank(jobs.size());
        for (size_t i = 0; i < jobs.size(); ++i) {
            size_t g = std::find(order.begin(), order.end(), jobs[i].scene) - order.begin();
            if (g == order.size()) order.push_back(jobs[i].scene);
            rank[i] = {g, i};
        }
        std::sort(rank.begin(), rank.end());
        std::vector<HostJob> sorted;
        sorted.reserve(jobs.size());
        for (const auto& gi : rank) sorted.push_back(jobs[gi.second]);
        jobs.swap(sorted);
    }
    for (size_t i = 0; i < jobs.size(); ++i)
        for (size_t k = 0; k < i; ++k)
            if (jobs[i].film && jobs[i].film == jobs[k].film)
                return fail(CN_E_INVALID, "two jobs of one wave fold into the same film");
    uint64_t slots = 0, pixels = 0;
    for (const HostJob& j : jobs) {
        slots += uint64_t(j.width) * j.height * j.samples;
        pixels += uint64_t(j.width) * j.height;
    }
    if (slots >= (1ull << 32)) return fail(CN_E_INVALID, "wave of %llu slots is too large", (unsigned long long)slots);
    PathPool pool;
    WaveTargets tg{};
    uint32_t capacity = r->capacity;
    if (arena) {
        if (slots > arena->slots)
            return fail(CN_E_INVALID, "a wave of %llu slots exceeds the arena's %zu", (unsigned long long)slots,
                        arena->slots);
        pool = arena->pool;
        pool.stack_overflow = r->pool.stack_overflow;
        pool.traversal_counts = r->pool.traversal_counts;
        capacity = std::min(capacity, arena->capacity);
        tg.radiance = arena->targets.radiance;
        if (aov) {
            tg.albedo = arena->targets.albedo;
            tg.normal = arena->targets.normal;
            tg.depth = arena->targets.depth;
            tg.guide = arena->targets.guide;
        }
    } else {
        if (int e = ensure_slots(r, size_t(slots), aov)) return e;
        if (int e = ensure_pool(r, size_t(slots))) return e;
        pool = r->pool;
        tg.radiance = r->radiance.ptr;
        if (aov) {
            tg.albedo = r->albedo.ptr;
            tg.normal = r->normal.ptr;
            tg.depth = r->depth.ptr;
            tg.guide = r->guide.ptr;
        }
    }

    // A job table from the ring: wait until the wave that last used it is done.
    const uint32_t e = r->ring_next;
    r->ring_next = (r->ring_next + 1) % JOB_RING;
    if (r->ring_used[e]) CN_HIP(hipEventSynchronize(r->ring_event[e]));
    JobRecord* host = r->staging + size_t(e) * MAX_JOBS;
    // With graphs, every wave reads its job table from one fixed device
    // table (stream order keeps a wave's table until its kernels are done),
    // so a captured launch sequence replays for any wave of the same shape.
    // (A wave that waits on another stream's event is launched directly.)
    const bool graph = r->graphs && !r->profiler && !r->counting && after_count == 0;
    JobRecord* dev = r->jobs.ptr + (graph ? 0 : size_t(e) * MAX_JOBS);
    uint32_t slot_base = 0, pixel_base = 0;
    for (size_t i = 0; i < jobs.size(); ++i) {
        const HostJob& j = jobs[i];
        JobRecord& rec = host[i];
        std::memset(&rec, 0, sizeof(rec));
        rec.cam = j.cam;
        rec.width = j.width;
        rec.height = j.height;
        rec.pixels = j.width * j.height;
        rec.slot_base = slot_base;
        rec.pixel_base = pixel_base;
        rec.first_sample = j.first_sample;
        rec.samples = j.samples;
        rec.flags = (j.film && j.first_sample == 0) ? JOB_RESET : 0u;
        rec.max_bounces = j.max_bounces;
        if (j.film) {
            rec.beauty_sum = j.film->beauty_sum.ptr;
            rec.albedo_sum = j.film->albedo_sum.ptr;
            rec.normal_sum = j.film->normal_sum.ptr;
            rec.depth_sum = j.film->depth_sum.ptr;
            rec.moment_sum = j.moments ? j.film->moment_sum.ptr : nullptr;
        }
        slot_base += rec.pixels * rec.samples;
        pixel_base += rec.pixels;
    }
    CN_HIP(hipMemcpyAsync(dev, host, jobs.size() * sizeof(JobRecord), hipMemcpyHostToDevice, r->stream));

    bool films = false;
    for (const HostJob& j : jobs) films = films || j.film != nullptr;

    // The wave's launch plan: per scene group, pool-sized ranges of
    // (base, count, bounce rounds).
    struct Range {
        size_t g0, g1;
        uint32_t base, count, rounds;
    };
    std::vector<Range> plan;
    size_t g0 = 0;
    while (g0 < jobs.size()) {
        size_t g1 = g0;
        while (g1 < jobs.size() && jobs[g1].scene == jobs[g0].scene) ++g1;
        uint32_t rounds = 0;  // bounce rounds: the group's deepest cap
        for (size_t k = g0; k < g1; ++k) rounds = std::max(rounds, jobs[k].max_bounces);
        const uint32_t s0 = host[g0].slot_base;
        const uint32_t s1 = host[g1 - 1].slot_base + host[g1 - 1].pixels * host[g1 - 1].samples;
        for (uint64_t base = s0; base < s1; base += capacity)
            plan.push_back({g0, g1, uint32_t(base), uint32_t(std::min<uint64_t>(capacity, s1 - base)), rounds});
        g0 = g1;
    }
    r->last_ranges = uint32_t(plan.size());

    const uint32_t per = counters_per_range(r->max_bounces);
    // The tail launch takes over after this bounce's trace (none: every
    // bounce as wavefront stages). Counting waves keep every bounce's queues.
    const uint32_t tail_bounce = (r->shape.tail_bounce == 0 || r->counting) ? UINT32_MAX : r->shape.tail_bounce;
    auto launch_all = [&]() -> int {
        for (size_t k = 0; k < plan.size(); ++k) {
            const Range& rg = plan[k];
            // Ranges run in stream order and raygen clears its counters, so
            // counter sets are reused past MAX_RANGES (only the per-wave
            // diagnostics need them distinct).
            uint32_t* counters = r->counters.ptr + (k % MAX_RANGES) * per;
            CN_HIP(launch_raygen(r->stream, counters, counter_lines_per_range(r->max_bounces), r->profiler.get()));
            for (uint32_t b = 0; b < rg.rounds; ++b) {
                const bool tail = b == tail_bounce;
                CN_HIP(launch_bounce(r->stream, r->shape, jobs[rg.g0].scene->view, pool, tg, counters, dev + rg.g0,
                                     uint32_t(rg.g1 - rg.g0), rg.base, rg.count, b, r->light_sampling, r->counting,
                                     tail, counters + tail_claim_index(r->max_bounces), r->ctx->compute_units,
                                     r->profiler.get()));
                if (tail) break;
            }
        }
        if (films) {
            for (uint32_t k = 0; k < after_count; ++k) CN_HIP(hipStreamWaitEvent(r->stream, after[k], 0));
            CN_HIP(launch_accumulate(r->stream, dev, uint32_t(jobs.size()), tg, uint32_t(pixels), r->profiler.get()));
        }
        return CN_OK;
    };

    if (!graph) {
        if (int err = launch_all()) return err;
    } else {
        // Everything that reaches a kernel argument, except the job table's
        // contents (read from device memory at run time).
        std::vector<uint64_t> key = {uint64_t(jobs.size()), pixels, uint64_t(films), uint64_t(aov),
                                     uint64_t(uintptr_t(tg.radiance)), uint64_t(uintptr_t(pool.origin)),
                                     uint64_t(uintptr_t(pool.stack_overflow)), r->shape.trace_blocks,
                                     r->shape.shade_variant, r->shape.shade_blocks, r->shape.light_blocks,
                                     r->shape.bvh_width, r->shape.tail_bounce, r->light_sampling};
        for (const Range& rg : plan) {
            key.insert(key.end(), {jobs[rg.g0].scene->serial, uint64_t(rg.g0), uint64_t(rg.g1), rg.base, rg.count,
                                   rg.rounds});
        }
        WaveGraph* hit = nullptr;
        for (WaveGraph& g : r->graph_cache)
            if (g.key == key) hit = &g;
        if (!hit) {
            if (r->graph_cache.size() >= 32) {
                auto lru = std::min_element(r->graph_cache.begin(), r->graph_cache.end(),
                                            [](const WaveGraph& x, const WaveGraph& y) { return x.last_use < y.last_use; });
                (void)hipGraphExecDestroy(lru->exec);
                r->graph_cache.erase(lru);
            }
            hipGraph_t g = nullptr;
            CN_HIP(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
            const int err = launch_all();
            const hipError_t end = hipStreamEndCapture(r->stream, &g);
            if (err) {
                if (g) (void)hipGraphDestroy(g);
                return err;
            }
            CN_HIP(end);
            hipGraphExec_t exec = nullptr;
            const hipError_t inst = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            CN_HIP(inst);
            r->graph_cache.push_back({std::move(key), exec, 0});
            hit = &r->graph_cache.back();
            r->graph_captures += 1;
        } else {
            r->graph_hits += 1;
        }
        hit->last_use = ++r->graph_clock;
        CN_HIP(hipGraphLaunch(hit->exec, r->stream));
    }
    CN_HIP(hipEventRecord(r->ring_event[e], r->stream));
    r->ring_used[e] = true;
    return CN_OK;
}

int scene_camera(cn_scene* s, const cn_camera& camera, uint32_t width, uint32_t height, CameraView* out) {
    if (!camera_view(camera, float(width) / float(height), out)) return fail(CN_E_INVALID, "degenerate camera");
    return CN_OK;
}

// Samples per wave for a view: as many as fill the path pool.
uint32_t samples_per_wave(const cn_renderer* r, uint32_t pixels, uint32_t wanted) {
    const uint32_t fit = std::max<uint32_t>(1, r->capacity / std::max<uint32_t>(1, pixels));
    return std::min({wanted, fit, 65535u});
}

int film_job(cn_renderer* r, cn_scene* s, cn_film* f, const cn_camera& camera, uint32_t samples, HostJob* job) {
    job->scene = s;
    job->width = f->width;
    job->height = f->height;
    job->first_sample = f->samples;
    job->samples = samples;
    job->film = f;
    return scene_camera(s, camera, f->width, f->height, &job->cam);
}

}  // namespace cn

extern "C" {

int cn_renderer_create(cn_context* ctx, uint32_t capacity, uint32_t max_bounces, uint32_t light_sampling,
                       cn_renderer** out) {
    if (!ctx || !out) return fail(CN_E_INVALID, "cn_renderer_create: null argument");
    if (capacity == 0) capacity = CN_POOL_CAPACITY;
    if (capacity > CN_POOL_CAPACITY_MAX)
        return fail(CN_E_INVALID, "capacity %u above the %u-path maximum", capacity, CN_POOL_CAPACITY_MAX);
    if (max_bounces == 0 || max_bounces > 255) return fail(CN_E_INVALID, "max_bounces %u outside 1..255", max_bounces);
    if (light_sampling > CN_LIGHT_SAMPLING_NEE) return fail(CN_E_INVALID, "light_sampling %u", light_sampling);
    CN_HIP(hipSetDevice(ctx->device));
    auto r = std::make_unique<cn_renderer>();
    r->ctx = ctx;
    r->capacity = capacity;
    r->max_bounces = max_bounces;
    r->light_sampling = light_sampling;
    const uint32_t cus = std::max(1u, ctx->compute_units);
    // Defaults from the autotuner's sweeps on the chart and heavy scenes;
    // cn_renderer_set_shape overrides them.
    r->shape.trace_blocks = cus * 64;
    r->shape.shade_variant = SHADE_256_GLOBAL_SORTED;
    r->shape.shade_blocks = cus * 16;
    r->shape.light_blocks = cus * 4;
    r->shape.bvh_width = 4;
    r->shape.tail_bounce = CN_DEFAULT_TAIL_BOUNCE;
    CN_HIP(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    CN_HIP(r->spill.alloc(size_t(r->shape.trace_blocks) * TRAVERSAL_BLOCK * STACK_SPILL));
    CN_HIP(r->traversal.alloc(4 * TRAVERSAL_ROWS));
    CN_HIP(hipMemset(r->traversal.ptr, 0, 4 * TRAVERSAL_ROWS * sizeof(unsigned long long)));
    CN_HIP(r->counters.alloc(size_t(MAX_RANGES) * counters_per_range(max_bounces)));
    CN_HIP(hipMemset(r->counters.ptr, 0, r->counters.count * 4));
    CN_HIP(r->jobs.alloc(size_t(JOB_RING) * MAX_JOBS));
    CN_HIP(hipHostMalloc(reinterpret_cast<void**>(&r->staging), sizeof(JobRecord) * JOB_RING * MAX_JOBS,
                         hipHostMallocDefault));
    for (uint32_t i = 0; i < JOB_RING; ++i) CN_HIP(hipEventCreateWithFlags(&r->ring_event[i], hipEventDisableTiming));
    r->pool.stack_overflow = r->spill.ptr;
    r->pool.traversal_counts = r->traversal.ptr;
    *out = r.release();
    return CN_OK;
}

int cn_renderer_destroy(cn_renderer* r) {
    if (!r) return fail(CN_E_INVALID, "cn_renderer_destroy: null renderer");
    (void)hipSetDevice(r->ctx->device);
    if (r->stream) {
        (void)hipStreamSynchronize(r->stream);
        (void)hipStreamDestroy(r->stream);
    }
    for (uint32_t i = 0; i < JOB_RING; ++i)
        if (r->ring_event[i]) (void)hipEventDestroy(r->ring_event[i]);
    if (r->staging) (void)hipHostFree(r->staging);
    delete r;
    return CN_OK;
}

int cn_renderer_reserve(cn_renderer* r, uint32_t paths, int aov) {
    if (!r) return fail(CN_E_INVALID, "cn_renderer_reserve: null renderer");
    CN_HIP(hipSetDevice(r->ctx->device));
    const size_t n = std::min<size_t>(paths, r->capacity);
    if (int e = ensure_slots(r, n, aov != 0)) return e;
    return ensure_pool(r, n);
}

int cn_renderer_get_shape(cn_renderer* r, cn_launch_shape* out) {
    if (!r || !out) return fail(CN_E_INVALID, "cn_renderer_get_shape: null argument");
    out->trace_blocks = r->shape.trace_blocks;
    out->shade_variant = r->shape.shade_variant;
    out->shade_blocks = r->shape.shade_blocks;
    out->light_blocks = r->shape.light_blocks;
    out->bvh_width = r->shape.bvh_width;
    out->tail_bounce = r->shape.tail_bounce;
    return CN_OK;
}

int cn_renderer_set_shape(cn_renderer* r, const cn_launch_shape* shape) {
    if (!r || !shape) return fail(CN_E_INVALID, "cn_renderer_set_shape: null argument");
    auto in_range = [](uint32_t v) { return v >= 1 && v <= 65536; };
    if (!in_range(shape->trace_blocks) || !in_range(shape->shade_blocks) || !in_range(shape->light_blocks))
        return fail(CN_E_INVALID, "launch shape block counts must lie in 1..65536");
    if (shape->shade_variant >= SHADE_VARIANTS)
        return fail(CN_E_INVALID, "unknown shade variant %u", shape->shade_variant);
    if (shape->bvh_width != 4 && shape->bvh_width != 8)
        return fail(CN_E_INVALID, "BVH width %u: 4 or 8", shape->bvh_width);
    if (shape->tail_bounce > 255) return fail(CN_E_INVALID, "tail bounce %u outside 0..255", shape->tail_bounce);
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    // The traversal stack's spill area is per lane of the trace grid.
    const size_t spill = size_t(shape->trace_blocks) * TRAVERSAL_BLOCK * STACK_SPILL;
    if (r->spill.count < spill) {
        r->drop_graphs();
        CN_HIP(r->spill.alloc(spill));
        r->pool.stack_overflow = r->spill.ptr;
    }
    r->shape.trace_blocks = shape->trace_blocks;
    r->shape.shade_variant = shape->shade_variant;
    r->shape.shade_blocks = shape->shade_blocks;
    r->shape.light_blocks = shape->light_blocks;
    r->shape.bvh_width = shape->bvh_width;
    r->shape.tail_bounce = shape->tail_bounce;
    return CN_OK;
}

int cn_renderer_set_graphs(cn_renderer* r, int enabled) {
    if (!r) return fail(CN_E_INVALID, "cn_renderer_set_graphs: null renderer");
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    r->graphs = enabled != 0;
    if (!r->graphs) r->drop_graphs();
    return CN_OK;
}

int cn_renderer_graph_stats(cn_renderer* r, uint64_t* captures, uint64_t* replays) {
    if (!r || !captures || !replays) return fail(CN_E_INVALID, "cn_renderer_graph_stats: null argument");
    *captures = r->graph_captures;
    *replays = r->graph_hits;
    return CN_OK;
}

int cn_renderer_set_counters(cn_renderer* r, int enabled) {
    if (!r) return fail(CN_E_INVALID, "cn_renderer_set_counters: null renderer");
    r->counting = enabled ? 1 : 0;
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    CN_HIP(hipMemset(r->traversal.ptr, 0, 4 * TRAVERSAL_ROWS * sizeof(unsigned long long)));
    return CN_OK;
}

int cn_renderer_set_profiling(cn_renderer* r, int enabled) {
    if (!r) return fail(CN_E_INVALID, "cn_renderer_set_profiling: null renderer");
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    if (enabled) {
        r->profiler = std::make_unique<EventProfiler>();
        r->profiler->stream = r->stream;
    } else {
        r->profiler.reset();
    }
    return CN_OK;
}

int cn_renderer_kernel_times(cn_renderer* r, cn_kernel_times* out) {
    if (!r || !out) return fail(CN_E_INVALID, "cn_renderer_kernel_times: null argument");
    if (!r->profiler) return fail(CN_E_INVALID, "profiling is off");
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    r->profiler->collect();
    static const char* names[K_KINDS] = {"raygen", "trace", "shade_miss", "shade_surface", "accumulate", "tail"};
    std::memset(out, 0, sizeof(*out));
    out->kernels = K_KINDS;
    for (int k = 0; k < K_KINDS; ++k) {
        std::snprintf(out->names[k], sizeof(out->names[k]), "%s", names[k]);
        out->total_ms[k] = r->profiler->total_ms[k];
        out->launches[k] = r->profiler->launches[k];
        r->profiler->total_ms[k] = 0.0;
        r->profiler->launches[k] = 0;
    }
    return CN_OK;
}

int cn_renderer_traversal_counts(cn_renderer* r, cn_traversal_counts* out) {
    if (!r || !out) return fail(CN_E_INVALID, "cn_renderer_traversal_counts: null argument");
    std::vector<unsigned long long> c(4 * TRAVERSAL_ROWS);
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    CN_HIP(hipMemcpy(c.data(), r->traversal.ptr, c.size() * sizeof(c[0]), hipMemcpyDeviceToHost));
    unsigned long long sum[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < c.size(); ++i) sum[i % 4] += c[i];
    out->ray_nodes = sum[0];
    out->ray_triangles = sum[1];
    out->shadow_nodes = sum[2];
    out->shadow_triangles = sum[3];
    return CN_OK;
}

int cn_renderer_traversal_counts_by_bounce(cn_renderer* r, uint64_t* out, uint32_t bounces) {
    if (!r || !out) return fail(CN_E_INVALID, "cn_renderer_traversal_counts_by_bounce: null argument");
    if (bounces > TRAVERSAL_ROWS) return fail(CN_E_INVALID, "at most %u bounces", TRAVERSAL_ROWS);
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    CN_HIP(hipMemcpy(out, r->traversal.ptr, size_t(4) * bounces * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return CN_OK;
}

int cn_renderer_wave_counts(cn_renderer* r, cn_wave_counts* out) {
    if (!r || !out) return fail(CN_E_INVALID, "cn_renderer_wave_counts: null argument");
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    if (r->last_ranges > MAX_RANGES)
        return fail(CN_E_INVALID, "the last wave walked %u pool ranges; counts cover at most %u", r->last_ranges,
                    MAX_RANGES);
    const uint32_t per = counters_per_range(r->max_bounces);
    std::vector<uint32_t> c(size_t(per) * r->last_ranges);
    CN_HIP(hipMemcpy(c.data(), r->counters.ptr, c.size() * 4, hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    out->bounces = r->max_bounces;
    for (uint32_t range = 0; range < r->last_ranges; ++range) {
        const uint32_t* rc = &c[size_t(range) * per];
        auto count = [&](uint32_t b, Queue q) {
            uint32_t sum = 0;
            for (uint32_t k = 0; k < QUEUE_SPLIT; ++k) sum += rc[counter_index(b, q, k)];
            return sum;
        };
        for (uint32_t b = 0; b < r->max_bounces && b < 256; ++b) {
            out->rays[b] += b == 0 ? count(0, Q_HIT) + count(0, Q_MISS) : count(b, Q_RAY);
            out->hits[b] += count(b, Q_HIT);
            out->misses[b] += count(b, Q_MISS);
            out->shadows[b] += count(b, Q_SHADOW);
        }
    }
    return CN_OK;
}

int cn_trace_aov(cn_renderer* r, cn_scene* s, float* radiance, float* albedo, float* normal, float* depth,
                 uint32_t width, uint32_t height, uint32_t sample) {
    if (!r || !s) return fail(CN_E_INVALID, "cn_trace: null renderer or scene");
    if (width == 0 || height == 0) return fail(CN_E_INVALID, "zero-sized trace target");
    CN_HIP(hipSetDevice(r->ctx->device));
    const size_t pixels = size_t(width) * height;
    const bool aov = albedo || normal || depth;
    HostJob job;
    job.scene = s;
    job.width = width;
    job.height = height;
    job.first_sample = sample;
    job.samples = 1;
    if (int e = scene_camera(s, s->camera, width, height, &job.cam)) return e;
    if (int e = enqueue_wave(r, {job}, aov)) return e;
    CN_HIP(hipStreamSynchronize(r->stream));
    // The slots hold the view in band order (band_pixel); the images are
    // row-major.
    auto fetch = [&](void* out, const void* slots, size_t bytes) -> int {
        std::vector<unsigned char> tmp(pixels * bytes);
        CN_HIP(hipMemcpy(tmp.data(), slots, tmp.size(), hipMemcpyDeviceToHost));
        unsigned char* dst = static_cast<unsigned char*>(out);
        for (uint32_t i = 0; i < pixels; ++i)
            std::memcpy(dst + size_t(band_pixel(i, width, height)) * bytes, tmp.data() + size_t(i) * bytes, bytes);
        return CN_OK;
    };
    if (radiance)
        if (int e = fetch(radiance, r->radiance.ptr, 16)) return e;
    if (albedo)
        if (int e = fetch(albedo, r->albedo.ptr, 16)) return e;
    if (normal)
        if (int e = fetch(normal, r->normal.ptr, 16)) return e;
    if (depth)
        if (int e = fetch(depth, r->depth.ptr, 4)) return e;
    return CN_OK;
}

int cn_trace(cn_renderer* r, cn_scene* s, float* radiance, uint32_t width, uint32_t height, uint32_t sample) {
    return cn_trace_aov(r, s, radiance, nullptr, nullptr, nullptr, width, height, sample);
}

int cn_first_hits(cn_renderer* r, cn_scene* s, uint32_t width, uint32_t height, uint32_t sample, uint32_t* instance,
                  uint32_t* primitive, float* bary) {
    if (!r || !s || !instance || !primitive || !bary) return fail(CN_E_INVALID, "cn_first_hits: null argument");
    if (width == 0 || height == 0) return fail(CN_E_INVALID, "zero-sized target");
    CN_HIP(hipSetDevice(r->ctx->device));
    CameraView cam;
    if (int e = scene_camera(s, s->camera, width, height, &cam)) return e;
    const size_t pixels = size_t(width) * height;
    DevBuf<uint32_t> di, dp;
    DevBuf<float> db;
    CN_HIP(di.alloc(pixels));
    CN_HIP(dp.alloc(pixels));
    CN_HIP(db.alloc(2 * pixels));
    CN_HIP(launch_first_hits(r->stream, r->shape, s->view, cam, r->pool, width, height, sample, di.ptr, dp.ptr, db.ptr));
    CN_HIP(hipStreamSynchronize(r->stream));
    CN_HIP(hipMemcpy(instance, di.ptr, pixels * 4, hipMemcpyDeviceToHost));
    CN_HIP(hipMemcpy(primitive, dp.ptr, pixels * 4, hipMemcpyDeviceToHost));
    CN_HIP(hipMemcpy(bary, db.ptr, pixels * 8, hipMemcpyDeviceToHost));
    return CN_OK;
}

int cn_film_create(cn_context* ctx, uint32_t width, uint32_t height, cn_film** out) {
    if (!ctx || !out) return fail(CN_E_INVALID, "cn_film_create: null argument");
    if (width == 0 || height == 0) return fail(CN_E_INVALID, "zero-sized film");
    CN_HIP(hipSetDevice(ctx->device));
    auto f = std::make_unique<cn_film>();
    f->ctx = ctx;
    f->width = width;
    f->height = height;
    const size_t n = size_t(width) * height;
    CN_HIP(f->beauty_sum.alloc(n));
    CN_HIP(f->albedo_sum.alloc(n));
    CN_HIP(f->normal_sum.alloc(n));
    CN_HIP(f->depth_sum.alloc(n));
    *out = f.release();
    return CN_OK;
}

int cn_film_destroy(cn_film* f) {
    if (!f) return fail(CN_E_INVALID, "cn_film_destroy: null film");
    (void)hipSetDevice(f->ctx->device);
    (void)hipDeviceSynchronize();
    delete f;
    return CN_OK;
}

int cn_film_reset(cn_film* f) {
    if (!f) return fail(CN_E_INVALID, "cn_film_reset: null film");
    f->samples = 0;
    return CN_OK;
}

int cn_film_samples(cn_film* f, uint32_t* samples) {
    if (!f || !samples) return fail(CN_E_INVALID, "cn_film_samples: null argument");
    *samples = f->samples;
    return CN_OK;
}

int cn_accumulate_async(cn_renderer* r, cn_scene* s, cn_film* f, uint32_t count) {
    if (!r || !s || !f) return fail(CN_E_INVALID, "cn_accumulate: null argument");
    CN_HIP(hipSetDevice(r->ctx->device));
    uint32_t left = count;
    while (left > 0) {
        const uint32_t n = samples_per_wave(r, f->width * f->height, left);
        HostJob job;
        if (int e = film_job(r, s, f, s->camera, n, &job)) return e;
        if (int e = enqueue_wave(r, {job}, true)) return e;
        f->samples += n;
        left -= n;
    }
    return CN_OK;
}

int cn_accumulate(cn_renderer* r, cn_scene* s, cn_film* f) {
    if (int e = cn_accumulate_async(r, s, f, 1)) return e;
    CN_HIP(hipStreamSynchronize(r->stream));
    return CN_OK;
}

int cn_render_batch(cn_renderer* r, const cn_view_job* views, uint32_t count) {
    if (!r || (!views && count)) return fail(CN_E_INVALID, "cn_render_batch: null argument");
    CN_HIP(hipSetDevice(r->ctx->device));
    std::vector<uint32_t> left(count);
    for (uint32_t i = 0; i < count; ++i) {
        if (!views[i].scene || !views[i].film) return fail(CN_E_INVALID, "view %u lacks a scene or film", i);
        left[i] = views[i].samples;
    }
    // Waves interleave the views at sample granularity: every unfinished
    // view gets its next sample while the pool has room, then leftover room
    // goes to more samples of the same views, in order.
    while (true) {
        std::vector<HostJob> wave;
        std::vector<uint32_t> who;
        uint64_t budget = r->capacity;
        for (uint32_t i = 0; i < count && wave.size() < MAX_JOBS; ++i) {
            if (!left[i]) continue;
            const uint64_t px = uint64_t(views[i].film->width) * views[i].film->height;
            if (px > budget && !wave.empty()) continue;
            HostJob job;
            if (int e = film_job(r, views[i].scene, views[i].film, views[i].camera, 1, &job)) return e;
            job.max_bounces = views[i].max_bounces;
            wave.push_back(job);
            who.push_back(i);
            budget = px > budget ? 0 : budget - px;
        }
        if (wave.empty()) break;
        for (size_t k = 0; k < wave.size(); ++k) {
            const uint32_t i = who[k];
            const uint64_t px = uint64_t(wave[k].width) * wave[k].height;
            const uint64_t extra = std::min<uint64_t>({uint64_t(left[i] - 1), budget / px, 65534u});
            wave[k].samples += uint32_t(extra);
            budget -= extra * px;
        }
        if (int e = enqueue_wave(r, wave, true)) return e;
        for (size_t k = 0; k < wave.size(); ++k) {
            left[who[k]] -= wave[k].samples;
            views[who[k]].film->samples += wave[k].samples;
        }
    }
    CN_HIP(hipStreamSynchronize(r->stream));
    return CN_OK;
}

int cn_renderer_synchronize(cn_renderer* r) {
    if (!r) return fail(CN_E_INVALID, "cn_renderer_synchronize: null renderer");
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    return CN_OK;
}

int cn_film_sums(cn_film* f, float* beauty, float* albedo, float* normal, float* depth) {
    if (!f) return fail(CN_E_INVALID, "cn_film_sums: null film");
    CN_HIP(hipSetDevice(f->ctx->device));
    CN_HIP(hipDeviceSynchronize());
    const size_t n = size_t(f->width) * f->height;
    if (beauty) CN_HIP(hipMemcpy(beauty, f->beauty_sum.ptr, n * 16, hipMemcpyDeviceToHost));
    if (albedo) CN_HIP(hipMemcpy(albedo, f->albedo_sum.ptr, n * 16, hipMemcpyDeviceToHost));
    if (normal) CN_HIP(hipMemcpy(normal, f->normal_sum.ptr, n * 16, hipMemcpyDeviceToHost));
    if (depth) CN_HIP(hipMemcpy(depth, f->depth_sum.ptr, n * 4, hipMemcpyDeviceToHost));
    return CN_OK;
}

int cn_film_averages(cn_film* f, float* beauty, float* albedo, float* normal, float* depth) {
    if (!f) return fail(CN_E_INVALID, "cn_film_averages: null film");
    if (f->samples == 0) return fail(CN_E_INVALID, "averaging an empty film");
    const size_t n = size_t(f->width) * f->height;
    if (int e = cn_film_sums(f, beauty, albedo, normal, depth)) return e;
    // render/film.rs:194-197: sum / samples on the host.
    const float samples = float(f->samples);
    auto divide = [&](float* a, size_t count) {
        if (a)
            for (size_t i = 0; i < count; ++i) a[i] = a[i] / samples;
    };
    divide(beauty, 4 * n);
    divide(albedo, 4 * n);
    divide(normal, 4 * n);
    divide(depth, n);
    return CN_OK;
}

int cn_film_set_sample(cn_film* f, const float* beauty, const float* albedo, const float* normal,
                       const float* depth) {
    if (!f) return fail(CN_E_INVALID, "cn_film_set_sample: null film");
    CN_HIP(hipSetDevice(f->ctx->device));
    const size_t n = size_t(f->width) * f->height;
    if (!f->beauty.ptr) {
        CN_HIP(f->beauty.alloc(n));
        CN_HIP(f->albedo.alloc(n));
        CN_HIP(f->normal.alloc(n));
        CN_HIP(f->depth.alloc(n));
        CN_HIP(hipMemset(f->beauty.ptr, 0, n * 16));
        CN_HIP(hipMemset(f->albedo.ptr, 0, n * 16));
        CN_HIP(hipMemset(f->normal.ptr, 0, n * 16));
        CN_HIP(hipMemset(f->depth.ptr, 0, n * 4));
    }
    if (beauty) CN_HIP(hipMemcpy(f->beauty.ptr, beauty, n * 16, hipMemcpyHostToDevice));
    if (albedo) CN_HIP(hipMemcpy(f->albedo.ptr, albedo, n * 16, hipMemcpyHostToDevice));
    if (normal) CN_HIP(hipMemcpy(f->normal.ptr, normal, n * 16, hipMemcpyHostToDevice));
    if (depth) CN_HIP(hipMemcpy(f->depth.ptr, depth, n * 4, hipMemcpyHostToDevice));
    return CN_OK;
}

int cn_film_accumulate_sample(cn_renderer* r, cn_film* f) {
    if (!r || !f) return fail(CN_E_INVALID, "cn_film_accumulate_sample: null argument");
    if (!f->beauty.ptr) return fail(CN_E_INVALID, "no sample was set on this film");
    CN_HIP(hipSetDevice(r->ctx->device));
    const uint32_t e = r->ring_next;
    r->ring_next = (r->ring_next + 1) % JOB_RING;
    if (r->ring_used[e]) CN_HIP(hipEventSynchronize(r->ring_event[e]));
    JobRecord* host = r->staging + size_t(e) * MAX_JOBS;
    std::memset(host, 0, sizeof(JobRecord));
    host->width = f->width;
    host->height = f->height;
    host->pixels = f->width * f->height;
    host->samples = 1;
    host->first_sample = f->samples;
    host->flags = f->samples == 0 ? JOB_RESET : 0u;
    host->beauty_sum = f->beauty_sum.ptr;
    host->albedo_sum = f->albedo_sum.ptr;
    host->normal_sum = f->normal_sum.ptr;
    host->depth_sum = f->depth_sum.ptr;
    JobRecord* dev = r->jobs.ptr + size_t(e) * MAX_JOBS;
    CN_HIP(hipMemcpyAsync(dev, host, sizeof(JobRecord), hipMemcpyHostToDevice, r->stream));
    WaveTargets tg{f->beauty.ptr, f->albedo.ptr, f->normal.ptr, f->depth.ptr, nullptr};
    CN_HIP(launch_accumulate(r->stream, dev, 1, tg, host->pixels, nullptr));
    CN_HIP(hipEventRecord(r->ring_event[e], r->stream));
    r->ring_used[e] = true;
    f->samples += 1;
    CN_HIP(hipStreamSynchronize(r->stream));
    return CN_OK;
}

int cn_resolve(cn_renderer* r, cn_film* f, float* beauty, float* albedo, float* normal, float* depth) {
    if (!r || !f) return fail(CN_E_INVALID, "cn_resolve: null argument");
    if (f->samples == 0) return fail(CN_E_INVALID, "resolving an empty film");
    CN_HIP(hipSetDevice(r->ctx->device));
    const size_t n = size_t(f->width) * f->height;
    if (!f->out_beauty.ptr) {
        CN_HIP(f->out_beauty.alloc(n));
        CN_HIP(f->out_albedo.alloc(n));
        CN_HIP(f->out_normal.alloc(n));
        CN_HIP(f->out_depth.alloc(n));
    }
    FilmView v{f->out_beauty.ptr, f->beauty_sum.ptr, f->out_albedo.ptr, f->albedo_sum.ptr,
               f->out_normal.ptr, f->normal_sum.ptr, f->out_depth.ptr, f->depth_sum.ptr};
    CN_HIP(launch_resolve(r->stream, v, uint32_t(n), float(f->samples)));
    CN_HIP(hipStreamSynchronize(r->stream));
    if (beauty) CN_HIP(hipMemcpy(beauty, f->out_beauty.ptr, n * 16, hipMemcpyDeviceToHost));
    if (albedo) CN_HIP(hipMemcpy(albedo, f->out_albedo.ptr, n * 16, hipMemcpyDeviceToHost));
    if (normal) CN_HIP(hipMemcpy(normal, f->out_normal.ptr, n * 16, hipMemcpyDeviceToHost));
    if (depth) CN_HIP(hipMemcpy(depth, f->out_depth.ptr, n * 4, hipMemcpyDeviceToHost));
    return CN_OK;
}

int cn_sample_sequence(cn_context* ctx, uint32_t pixel, uint32_t dimension, uint32_t count, float* points2d,
                       float* values1d) {
    if (!ctx || !points2d || !values1d) return fail(CN_E_INVALID, "cn_sample_sequence: null argument");
    if (count == 0) return CN_OK;
    CN_HIP(hipSetDevice(ctx->device));
    DevBuf<float> p, v;
    CN_HIP(p.alloc(2 * size_t(count)));
    CN_HIP(v.alloc(count));
    CN_HIP(launch_sample_sequence(nullptr, pixel, dimension, count, p.ptr, v.ptr));
    CN_HIP(hipDeviceSynchronize());
    CN_HIP(hipMemcpy(points2d, p.ptr, 8 * size_t(count), hipMemcpyDeviceToHost));
    CN_HIP(hipMemcpy(values1d, v.ptr, 4 * size_t(count), hipMemcpyDeviceToHost));
    return CN_OK;
}

}  // extern "C"
