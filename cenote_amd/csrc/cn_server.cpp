// cn_server.cpp — the serving loop: progressive render requests admitted
// by an HBM budget and interleaved at sample granularity into shared waves
// (continuous batching). The reference's nearest relative is the
// progressive Session (render/session.rs:480-489), which renders one view;
// this loop generalises it to many concurrent requests.
//
// The server runs its waves on LANES lanes — renderers of its own, each with
// its stream and path pool, one wave in flight each. Lane 0 carries first
// frames: a wave of one sample for every admitted request that has none
// yet, at most wave_paths paths, launched as soon as the lane is free so
// that a new request's first frame never queues behind bulk work. The other
// lanes carry bulk waves — 4 x wave_paths paths while requests wait for a
// first frame, 8 x otherwise — of one more sample for every request whose
// first sample is out, then leftover room as extra samples in admission
// order. Bulk waves are large, so their sparse
// deep bounces are a small share of them, and the lanes overlap one wave's
// deep bounces with the others' full ones. A wave's accumulate waits for the
// earlier waves that fold into any of its films (events across lanes), so
// every film sees its samples in index order and a result is bitwise the
// same whether the request is served alone or in any mix. Waves retire as
// they complete, in any order.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "cn_handles.h"

using namespace cn;

namespace {

constexpr uint32_t NOISE_BLOCKS = 64;  // partial sums per noise estimate
constexpr uint32_t LANES = 3;          // lane 0: first frames; lanes 1 and 2: bulk waves
// A bulk wave holds up to BULK_FACTOR x wave_paths paths while requests wait
// for their first frame (smaller bulk waves leave the first-frame lane more
// of the GPU), and up to BULK_FACTOR_IDLE x wave_paths otherwise (larger
// waves spend less of their time in sparse deep bounces).
constexpr uint32_t BULK_FACTOR = 4;
constexpr uint32_t BULK_FACTOR_IDLE = 8;

size_t film_bytes(uint32_t width, uint32_t height, bool noise) {
    // beauty, albedo, normal sums (16 B each) + depth (4 B) [+ moments (4 B)]
    return size_t(width) * height * (52 + (noise ? 4 : 0));
}

struct Request {
    uint64_t id = 0;
    cn_request_desc desc{};
    uint32_t state = CN_REQUEST_QUEUED;
    double arrival = 0.0, admitted = -1.0, first_frame = -1.0, done = -1.0;
    cn_film* film = nullptr;
    uint32_t scheduled = 0;  // samples enqueued
    uint32_t completed = 0;  // samples whose waves have retired
    bool converged = false;  // the noise target was met
    double noise = -1.0;     // last relMSE estimate
    size_t bytes = 0;
    uint64_t last_wave = 0;  // serial of the last wave enqueued for this film (0: none)
};

struct WavePart {
    uint64_t id;
    uint32_t samples_after;
    int noise_slot;  // index into the wave's noise partials, or -1
};

struct InFlight {
    uint64_t serial = 0;
    uint32_t lane = 0;
    hipEvent_t event = nullptr;
    uint32_t ring = 0;  // noise-partials ring entry
    std::vector<WavePart> parts;
    uint64_t paths = 0;
};

}  // namespace

struct cn_server {
    cn_renderer* r = nullptr;           // the caller's renderer: depth, light sampling, launch shape
    std::vector<cn_renderer*> lanes;    // the server's own renderers, one wave in flight each
    bool lane_busy[LANES] = {};
    uint64_t wave_serial = 0;
    uint64_t budget = 0;
    uint32_t wave_paths = 0;       // first-frame waves
    uint32_t bulk_paths = 0;       // bulk waves while first frames are pending
    uint32_t bulk_paths_idle = 0;  // bulk waves otherwise
    std::chrono::steady_clock::time_point t0;
    std::map<uint64_t, Request> requests;
    std::deque<uint64_t> queue;    // waiting for admission, FIFO
    std::vector<uint64_t> active;  // admitted, in admission order
    uint64_t used = 0;             // bytes of films held by requests
    // Released films kept for reuse (freeing device memory would stall the
    // pipeline), by size; they count against the budget until evicted.
    std::multimap<std::pair<uint32_t, uint32_t>, cn_film*> spare;
    uint64_t spare_bytes = 0;
    uint64_t next_id = 1;
    std::vector<InFlight> inflight;  // in enqueue order
    std::vector<hipEvent_t> free_events;
    // Noise partials: per lane, MAX_JOBS estimates of NOISE_BLOCKS partial sums.
    DevBuf<double> partials;
    double* host_partials = nullptr;  // pinned
    uint64_t progress = 0;  // admissions + waves enqueued + waves retired
    cn_server_stats stats{};

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// Noise partials: one entry of MAX_JOBS estimates per lane (a lane holds
// one wave in flight).
constexpr uint32_t NOISE_RING = LANES;

int admit(cn_server* sv) {
    while (!sv->queue.empty() && sv->active.size() < MAX_JOBS) {
        Request& q = sv->requests[sv->queue.front()];
        // Spares may be evicted to make room; held films may not.
        if (sv->used + q.bytes > sv->budget) break;
        cn_film* f = nullptr;
        const auto size = std::make_pair(q.desc.width, q.desc.height);
        auto hit = sv->spare.find(size);
        if (hit != sv->spare.end()) {
            f = hit->second;
            sv->spare.erase(hit);
            sv->spare_bytes -= film_bytes(size.first, size.second, f->moment_sum.ptr != nullptr);
            f->samples = 0;
        } else {
            // Evict spares of other sizes until the new film fits.
            while (!sv->spare.empty() && sv->used + sv->spare_bytes + q.bytes > sv->budget) {
                auto victim = sv->spare.begin();
                sv->spare_bytes -= film_bytes(victim->first.first, victim->first.second,
                                              victim->second->moment_sum.ptr != nullptr);
                delete victim->second;
                sv->spare.erase(victim);
            }
            if (int e = cn_film_create(sv->r->ctx, q.desc.width, q.desc.height, &f)) return e;
        }
        if (q.desc.rel_mse > 0.0f && !f->moment_sum.ptr) {
            const hipError_t he = f->moment_sum.alloc(size_t(q.desc.width) * q.desc.height);
            if (he != hipSuccess) {
                delete f;
                return fail(CN_E_ALLOC, "noise moments: %s", hipGetErrorString(he));
            }
        }
        q.film = f;
        q.state = CN_REQUEST_ACTIVE;
        q.admitted = sv->now();
        q.bytes = film_bytes(q.desc.width, q.desc.height, f->moment_sum.ptr != nullptr);
        sv->used += q.bytes;
        sv->active.push_back(q.id);
        sv->queue.pop_front();
        sv->progress += 1;
    }
    return CN_OK;
}

bool wants_more(const Request& q) { return !q.converged && q.scheduled < q.desc.spp; }

// Retires in-flight wave k, whose event has completed.
int retire(cn_server* sv, size_t k) {
    InFlight w = std::move(sv->inflight[k]);
    sv->inflight.erase(sv->inflight.begin() + ptrdiff_t(k));
    CN_HIP(hipEventSynchronize(w.event));
    sv->free_events.push_back(w.event);
    sv->lane_busy[w.lane] = false;
    sv->progress += 1;
    const double t = sv->now();
    sv->stats.waves += 1;
    sv->stats.paths += w.paths;
    const double* noise = sv->host_partials + size_t(w.ring) * MAX_JOBS * NOISE_BLOCKS;
    for (const WavePart& part : w.parts) {
        Request& q = sv->requests[part.id];
        // A film's waves complete in enqueue order (each accumulate waits
        // for the film's previous wave), so the count only grows.
        q.completed = std::max(q.completed, part.samples_after);
        if (q.first_frame < 0.0 && q.completed >= 1) q.first_frame = t;
        if (part.noise_slot >= 0) {
            double sum = 0.0;
            for (uint32_t b = 0; b < NOISE_BLOCKS; ++b) sum += noise[size_t(part.noise_slot) * NOISE_BLOCKS + b];
            q.noise = sum / (double(q.desc.width) * q.desc.height);
            if (q.noise <= q.desc.rel_mse) q.converged = true;
        }
    }
    // Requests whose every scheduled sample has retired and that want no
    // more are done; their films stay until the caller takes the result.
    for (auto it = sv->active.begin(); it != sv->active.end();) {
        Request& q = sv->requests[*it];
        if (q.completed == q.scheduled && !wants_more(q)) {
            q.state = CN_REQUEST_DONE;
            q.done = t;
            sv->stats.completed += 1;
            it = sv->active.erase(it);
        } else {
            ++it;
        }
    }
    return CN_OK;
}

// Retires every in-flight wave that has completed; *count gets how many.
int retire_ready(cn_server* sv, uint32_t* count) {
    *count = 0;
    for (size_t k = 0; k < sv->inflight.size();) {
        const hipError_t q = hipEventQuery(sv->inflight[k].event);
        if (q == hipErrorNotReady) {
            ++k;
            continue;
        }
        CN_HIP(q);
        if (int e = retire(sv, k)) return e;
        *count += 1;
    }
    return CN_OK;
}

// Builds and enqueues one wave on `lane`: with `first`, the first sample of
// every admitted request that has none (at most wave_paths paths, the first
// request always); otherwise one more sample for every request whose first
// sample is out, then leftover room as extra samples (at most bulk_paths).
// *enqueued tells whether there was anything to enqueue.
int enqueue(cn_server* sv, uint32_t lane, bool first, bool* enqueued) {
    *enqueued = false;
    cn_renderer* r = sv->lanes[lane];
    std::vector<HostJob> jobs;
    std::vector<uint64_t> ids;
    std::vector<uint32_t> counts;
    bool waiting = !sv->queue.empty();
    for (uint64_t id : sv->active) waiting = waiting || sv->requests[id].scheduled == 0;
    uint64_t room = first ? sv->wave_paths : (waiting ? sv->bulk_paths : sv->bulk_paths_idle);
    auto add = [&](Request& q, uint32_t n) -> int {
        HostJob job;
        if (int e = film_job(r, q.desc.scene, q.film, q.desc.camera, n, &job)) return e;
        job.first_sample = q.scheduled;
        job.max_bounces = q.desc.max_bounces;
        job.moments = q.desc.rel_mse > 0.0f;
        jobs.push_back(job);
        ids.push_back(q.id);
        counts.push_back(n);
        return CN_OK;
    };
    for (uint64_t id : sv->active) {
        Request& q = sv->requests[id];
        if (!wants_more(q) || jobs.size() >= MAX_JOBS) continue;
        if (first != (q.scheduled == 0)) continue;
        const uint64_t px = uint64_t(q.desc.width) * q.desc.height;
        if (px > room && !jobs.empty()) continue;
        if (int e = add(q, 1)) return e;
        room = px > room ? 0 : room - px;
    }
    if (jobs.empty()) return CN_OK;
    if (!first) {
        // Leftover room: more samples of the same requests, in order.
        for (size_t k = 0; k < jobs.size(); ++k) {
            Request& q = sv->requests[ids[k]];
            const uint64_t px = uint64_t(q.desc.width) * q.desc.height;
            const uint64_t extra = std::min<uint64_t>({uint64_t(q.desc.spp - q.scheduled - 1), room / px, 65534u});
            jobs[k].samples += uint32_t(extra);
            counts[k] += uint32_t(extra);
            room -= extra * px;
        }
    }
    // The accumulate waits for the in-flight waves that fold into any of
    // this wave's films.
    std::vector<hipEvent_t> after;
    for (uint64_t id : ids) {
        const uint64_t last = sv->requests[id].last_wave;
        for (const InFlight& w : sv->inflight)
            if (w.serial == last && std::find(after.begin(), after.end(), w.event) == after.end())
                after.push_back(w.event);
    }
    if (int e = enqueue_wave(r, jobs, true, nullptr, after.data(), uint32_t(after.size()))) return e;
    InFlight w;
    w.serial = ++sv->wave_serial;
    w.lane = lane;
    w.ring = lane;
    double* dev_partials = sv->partials.ptr + size_t(w.ring) * MAX_JOBS * NOISE_BLOCKS;
    double* host_partials = sv->host_partials + size_t(w.ring) * MAX_JOBS * NOISE_BLOCKS;
    int noise_slots = 0;
    for (size_t k = 0; k < jobs.size(); ++k) {
        Request& q = sv->requests[ids[k]];
        q.scheduled += counts[k];
        q.film->samples = q.scheduled;
        q.last_wave = w.serial;
        w.paths += uint64_t(jobs[k].width) * jobs[k].height * counts[k];
        int slot = -1;
        if (q.desc.rel_mse > 0.0f && q.scheduled >= 2) {
            slot = noise_slots++;
            CN_HIP(launch_noise(r->stream, q.film->beauty_sum.ptr, q.film->moment_sum.ptr,
                                q.desc.width * q.desc.height, float(q.scheduled),
                                dev_partials + size_t(slot) * NOISE_BLOCKS, NOISE_BLOCKS));
        }
        w.parts.push_back({q.id, q.scheduled, slot});
    }
    if (noise_slots > 0) {
        CN_HIP(hipMemcpyAsync(host_partials, dev_partials, sizeof(double) * NOISE_BLOCKS * noise_slots,
                              hipMemcpyDeviceToHost, r->stream));
    }
    if (sv->free_events.empty()) {
        hipEvent_t e;
        CN_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        sv->free_events.push_back(e);
    }
    w.event = sv->free_events.back();
    sv->free_events.pop_back();
    CN_HIP(hipEventRecord(w.event, r->stream));
    sv->lane_busy[lane] = true;
    sv->inflight.push_back(std::move(w));
    sv->progress += 1;
    *enqueued = true;
    return CN_OK;
}

}  // namespace

extern "C" {

int cn_server_create(cn_renderer* renderer, uint64_t hbm_budget_bytes, uint32_t wave_paths, cn_server** out) {
    if (!renderer || !out) return fail(CN_E_INVALID, "cn_server_create: null argument");
    if (hbm_budget_bytes == 0) return fail(CN_E_INVALID, "cn_server_create: zero HBM budget");
    CN_HIP(hipSetDevice(renderer->ctx->device));
    auto sv = std::make_unique<cn_server>();
    sv->r = renderer;
    sv->budget = hbm_budget_bytes;
    sv->wave_paths = std::min(wave_paths == 0 ? (1u << 24) : wave_paths, renderer->capacity);
    sv->bulk_paths = uint32_t(std::min<uint64_t>(uint64_t(sv->wave_paths) * BULK_FACTOR, renderer->capacity));
    sv->bulk_paths_idle =
        uint32_t(std::min<uint64_t>(uint64_t(sv->wave_paths) * BULK_FACTOR_IDLE, renderer->capacity));
    sv->t0 = std::chrono::steady_clock::now();
    // The lanes copy the caller's renderer's depth, light sampling and launch
    // shape; each holds a pool of its waves' size, allocated here.
    auto lanes_up = [&]() -> int {
        cn_launch_shape shape;
        if (int e = cn_renderer_get_shape(renderer, &shape)) return e;
        for (uint32_t k = 0; k < LANES; ++k) {
            const uint32_t paths = k == 0 ? sv->wave_paths : sv->bulk_paths_idle;
            const uint32_t capacity = paths;
            cn_renderer* lane = nullptr;
            if (int e = cn_renderer_create(renderer->ctx, capacity, renderer->max_bounces, renderer->light_sampling,
                                           &lane))
                return e;
            sv->lanes.push_back(lane);
            if (int e = cn_renderer_set_shape(lane, &shape)) return e;
            if (int e = cn_renderer_reserve(lane, paths, 1)) return e;
        }
        CN_HIP(sv->partials.alloc(size_t(NOISE_RING) * MAX_JOBS * NOISE_BLOCKS));
        CN_HIP(hipHostMalloc(reinterpret_cast<void**>(&sv->host_partials),
                             sizeof(double) * NOISE_RING * MAX_JOBS * NOISE_BLOCKS, hipHostMallocDefault));
        return CN_OK;
    };
    if (int e = lanes_up()) {
        for (cn_renderer* lane : sv->lanes) cn_renderer_destroy(lane);
        return e;
    }
    *out = sv.release();
    return CN_OK;
}

int cn_server_destroy(cn_server* sv) {
    if (!sv) return fail(CN_E_INVALID, "cn_server_destroy: null server");
    (void)hipSetDevice(sv->r->ctx->device);
    for (cn_renderer* lane : sv->lanes) (void)hipStreamSynchronize(lane->stream);
    for (auto& w : sv->inflight) (void)hipEventDestroy(w.event);
    for (hipEvent_t e : sv->free_events) (void)hipEventDestroy(e);
    for (auto& [id, q] : sv->requests) delete q.film;
    for (auto& [size, f] : sv->spare) delete f;
    if (sv->host_partials) (void)hipHostFree(sv->host_partials);
    for (cn_renderer* lane : sv->lanes) cn_renderer_destroy(lane);
    delete sv;
    return CN_OK;
}

int cn_server_clock(cn_server* sv, double* seconds) {
    if (!sv || !seconds) return fail(CN_E_INVALID, "cn_server_clock: null argument");
    *seconds = sv->now();
    return CN_OK;
}

int cn_server_submit(cn_server* sv, const cn_request_desc* desc, uint64_t* id) {
    if (!sv || !desc || !id) return fail(CN_E_INVALID, "cn_server_submit: null argument");
    if (!desc->scene) return fail(CN_E_INVALID, "request without a scene");
    if (desc->width == 0 || desc->height == 0) return fail(CN_E_INVALID, "zero-sized request");
    if (desc->spp == 0) return fail(CN_E_INVALID, "request for zero samples");
    if (desc->max_bounces > sv->r->max_bounces)
        return fail(CN_E_INVALID, "request depth %u exceeds the renderer's %u", desc->max_bounces,
                    sv->r->max_bounces);
    if (!(desc->rel_mse >= 0.0f)) return fail(CN_E_INVALID, "negative or NaN noise target");
    CameraView probe;
    if (!camera_view(desc->camera, float(desc->width) / float(desc->height), &probe))
        return fail(CN_E_INVALID, "degenerate camera");
    const size_t bytes = film_bytes(desc->width, desc->height, desc->rel_mse > 0.0f);
    if (bytes > sv->budget) return fail(CN_E_ALLOC, "request needs %zu bytes, over the whole budget", bytes);
    Request q;
    q.id = sv->next_id++;
    q.desc = *desc;
    q.arrival = desc->arrival_seconds >= 0.0 ? desc->arrival_seconds : sv->now();
    q.bytes = bytes;
    sv->requests[q.id] = q;
    sv->queue.push_back(q.id);
    sv->stats.submitted += 1;
    *id = q.id;
    return CN_OK;
}

int cn_server_step(cn_server* sv, uint32_t* busy) {
    if (!sv) return fail(CN_E_INVALID, "cn_server_step: null server");
    CN_HIP(hipSetDevice(sv->r->ctx->device));
    uint32_t retired = 0;
    if (int e = retire_ready(sv, &retired)) return e;
    if (int e = admit(sv)) return e;
    bool enqueued = false;
    for (uint32_t lane = 0; lane < LANES; ++lane) {
        if (sv->lane_busy[lane]) continue;
        bool more = false;
        if (int e = enqueue(sv, lane, lane == 0, &more)) return e;
        enqueued = enqueued || more;
    }
    // Nothing new and nothing retired: wait a little for the next wave to
    // complete, so callers can loop on step without spinning, while new
    // arrivals still reach the free first-frame lane within a millisecond.
    if (!enqueued && retired == 0 && !sv->inflight.empty()) {
        const auto until = std::chrono::steady_clock::now() + std::chrono::milliseconds(1);
        while (retired == 0 && std::chrono::steady_clock::now() < until) {
            std::this_thread::sleep_for(std::chrono::microseconds(50));
            if (int e = retire_ready(sv, &retired)) return e;
        }
    }
    if (busy) *busy = uint32_t(sv->queue.size() + sv->active.size() + sv->inflight.size());
    return CN_OK;
}

int cn_server_drain(cn_server* sv) {
    if (!sv) return fail(CN_E_INVALID, "cn_server_drain: null server");
    uint32_t busy = 1;
    while (busy) {
        const uint64_t before = sv->progress;
        if (int e = cn_server_step(sv, &busy)) return e;
        if (busy && sv->progress == before && sv->inflight.empty()) {
            // Only queued requests remain and none fits: finished requests'
            // films hold the budget until their results are taken.
            return fail(CN_E_ALLOC, "%u queued request(s) wait for budget held by finished requests' films",
                        uint32_t(sv->queue.size()));
        }
    }
    return CN_OK;
}

int cn_server_status(cn_server* sv, uint64_t id, cn_request_status* out) {
    if (!sv || !out) return fail(CN_E_INVALID, "cn_server_status: null argument");
    auto it = sv->requests.find(id);
    if (it == sv->requests.end()) return fail(CN_E_INVALID, "unknown request %llu", (unsigned long long)id);
    const Request& q = it->second;
    out->id = q.id;
    out->state = q.state;
    out->samples = q.completed;
    out->arrival_seconds = q.arrival;
    out->admitted_seconds = q.admitted;
    out->first_frame_seconds = q.first_frame;
    out->done_seconds = q.done;
    out->rel_mse_estimate = q.noise;
    return CN_OK;
}

int cn_server_stats_get(cn_server* sv, cn_server_stats* out) {
    if (!sv || !out) return fail(CN_E_INVALID, "cn_server_stats_get: null argument");
    *out = sv->stats;
    out->queued = uint32_t(sv->queue.size());
    out->active = uint32_t(sv->active.size());
    out->hbm_used_bytes = sv->used;
    return CN_OK;
}

int cn_server_result(cn_server* sv, uint64_t id, float* beauty, float* albedo, float* normal, float* depth) {
    if (!sv) return fail(CN_E_INVALID, "cn_server_result: null server");
    auto it = sv->requests.find(id);
    if (it == sv->requests.end()) return fail(CN_E_INVALID, "unknown request %llu", (unsigned long long)id);
    Request& q = it->second;
    if (q.state != CN_REQUEST_DONE) return fail(CN_E_INVALID, "request %llu is not done", (unsigned long long)id);
    if (!q.film) return fail(CN_E_INVALID, "request %llu's result was already taken", (unsigned long long)id);
    CN_HIP(hipSetDevice(sv->r->ctx->device));
    // Every wave that touched this film has retired (its event completed).
    const size_t n = size_t(q.desc.width) * q.desc.height;
    cn_film* f = q.film;
    if (beauty) CN_HIP(hipMemcpy(beauty, f->beauty_sum.ptr, n * 16, hipMemcpyDeviceToHost));
    if (albedo) CN_HIP(hipMemcpy(albedo, f->albedo_sum.ptr, n * 16, hipMemcpyDeviceToHost));
    if (normal) CN_HIP(hipMemcpy(normal, f->normal_sum.ptr, n * 16, hipMemcpyDeviceToHost));
    if (depth) CN_HIP(hipMemcpy(depth, f->depth_sum.ptr, n * 4, hipMemcpyDeviceToHost));
    const float samples = float(q.completed);
    auto divide = [&](float* a, size_t count) {
        if (a)
            for (size_t i = 0; i < count; ++i) a[i] = a[i] / samples;
    };
    divide(beauty, 4 * n);
    divide(albedo, 4 * n);
    divide(normal, 4 * n);
    divide(depth, n);
    sv->spare.emplace(std::make_pair(q.desc.width, q.desc.height), f);
    sv->spare_bytes += q.bytes;
    q.film = nullptr;
    q.state = CN_REQUEST_RELEASED;
    sv->used -= q.bytes;
    return CN_OK;
}

}  // extern "C"
