// cn_static.hip — the allocation-free variant of cn_accumulate_async.
//
// Callers that cannot allocate device memory (code under stream capture,
// tight serving loops, harnesses that forbid hipMalloc) still need the
// wave's working memory: the path pool, its queues and shadow records, and
// the per-slot targets. Here that memory is one file-scope device array,
// `scratch`, sized for waves of up to STATIC_PATHS paths and slots. Each
// device the library runs on holds its own copy until the process ends; its
// address on each device is taken once, when a context is created on it
// (static_arena_init), and every later use goes through that address.
//
// Nothing in `scratch` has to start at zero: raygen writes every path it
// starts, the bounce-0 kernels store each slot target's first value, and
// the queue counters live in the renderer. Every call on one device shares
// the array, so two calls whose waves overlap in time on one device (two
// renderers' streams running at once) race; calls on one renderer are
// ordered by its stream and do not.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "cn_handles.h"

using namespace cn;

namespace {

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// The arena's layout: the pool's per-path arrays, then the queues and shadow
// records (QUEUE_SPLIT sub-queues of sub_capacity_for(STATIC_PATHS) entries),
// then the slot targets.
struct Layout {
    size_t origin, direction, throughput, hit, key, state;
    size_t ray0, ray1, hits, misses, shadow_origin, shadow_dir, shadow_contrib, shadow_prim;
    size_t radiance, albedo, normal, guide, depth;
    size_t total;
};

constexpr Layout make_layout() {
    const size_t P = STATIC_PATHS;
    const size_t E = size_t(sub_capacity_for(STATIC_PATHS)) * QUEUE_SPLIT;
    const size_t S = STATIC_PATHS;
    Layout l{};
    size_t at = 0;
    auto take = [&at](size_t bytes) {
        const size_t here = at;
        at = align256(at + bytes);
        return here;
    };
    l.origin = take(P * 16);
    l.direction = take(P * 16);
    l.throughput = take(P * 16);
    l.hit = take(P * 16);
    l.key = take(P * 8);
    l.state = take(P * 4);
    l.ray0 = take(E * 4);
    l.ray1 = take(E * 4);
    l.hits = take(E * 4);
    l.misses = take(E * 4);
    l.shadow_origin = take(E * 16);
    l.shadow_dir = take(E * 16);
    l.shadow_contrib = take(E * 16);
    l.shadow_prim = take(E * 4);
    l.radiance = take(S * 16);
    l.albedo = take(S * 16);
    l.normal = take(S * 16);
    l.guide = take(S * 16);
    l.depth = take(S * 4);
    l.total = at;
    return l;
}

constexpr Layout LAYOUT = make_layout();
static_assert(LAYOUT.total == STATIC_SCRATCH_BYTES, "the header's stated arena size matches its layout");

}  // namespace

static __device__ __align__(256) unsigned char scratch[LAYOUT.total];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex arena_mutex;
void* arena_base[MAX_DEVICES] = {};

template <typename T>
T* at(void* base, size_t offset) {
    return reinterpret_cast<T*>(static_cast<unsigned char*>(base) + offset);
}

WaveArena carve(void* base) {
    WaveArena a{};
    PathPool& p = a.pool;
    p.origin = at<float4>(base, LAYOUT.origin);
    p.direction = at<float4>(base, LAYOUT.direction);
    p.throughput = at<float4>(base, LAYOUT.throughput);
    p.hit = at<uint4>(base, LAYOUT.hit);
    p.key = at<uint2>(base, LAYOUT.key);
    p.state = at<uint32_t>(base, LAYOUT.state);
    p.ray_queue[0] = at<uint32_t>(base, LAYOUT.ray0);
    p.ray_queue[1] = at<uint32_t>(base, LAYOUT.ray1);
    p.hit_queue = at<uint32_t>(base, LAYOUT.hits);
    p.miss_queue = at<uint32_t>(base, LAYOUT.misses);
    p.shadow_origin = at<float4>(base, LAYOUT.shadow_origin);
    p.shadow_dir = at<float4>(base, LAYOUT.shadow_dir);
    p.shadow_contrib = at<float4>(base, LAYOUT.shadow_contrib);
    p.shadow_prim = at<uint32_t>(base, LAYOUT.shadow_prim);
    p.sub_capacity = sub_capacity_for(STATIC_PATHS);
    a.targets.radiance = at<float4>(base, LAYOUT.radiance);
    a.targets.albedo = at<float4>(base, LAYOUT.albedo);
    a.targets.normal = at<float4>(base, LAYOUT.normal);
    a.targets.guide = at<float4>(base, LAYOUT.guide);
    a.targets.depth = at<float>(base, LAYOUT.depth);
    a.capacity = STATIC_PATHS;
    a.slots = STATIC_PATHS;
    return a;
}

}  // namespace

namespace cn {

// Takes the address of `scratch` on `device`, which must be current. Called
// by cn_context_create, once per device.
int static_arena_init(int device) {
    if (device < 0 || device >= MAX_DEVICES) return fail(CN_E_NO_DEVICE, "device %d beyond the arena table", device);
    std::lock_guard<std::mutex> lock(arena_mutex);
    if (arena_base[device]) return CN_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(CN_E_DEVICE_FAULT, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    arena_base[device] = ptr;
    return CN_OK;
}

}  // namespace cn

extern "C" int cn_accumulate_async_static(cn_renderer* r, cn_scene* s, cn_film* f, uint32_t count) {
    if (!r || !s || !f) return fail(CN_E_INVALID, "cn_accumulate_async_static: null argument");
    const uint64_t pixels = uint64_t(f->width) * f->height;
    if (pixels > STATIC_PATHS)
        return fail(CN_E_INVALID,
                    "cn_accumulate_async_static: a %ux%u film (%llu pixels) exceeds the static arena's %u-pixel "
                    "maximum; use cn_accumulate_async",
                    f->width, f->height, (unsigned long long)pixels, STATIC_PATHS);
    const int device = r->ctx->device;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(arena_mutex);
        if (device >= 0 && device < MAX_DEVICES) base = arena_base[device];
    }
    if (!base) return fail(CN_E_INVALID, "cn_accumulate_async_static: no arena on device %d", device);
    CN_HIP(hipSetDevice(device));
    const WaveArena arena = carve(base);
    // As cn_accumulate_async, with waves of at most STATIC_PATHS slots.
    const uint32_t fit = std::max<uint32_t>(1, uint32_t(STATIC_PATHS / pixels));
    uint32_t left = count;
    while (left > 0) {
        const uint32_t n = std::min({left, fit, 65535u});
        HostJob job;
        if (int e = film_job(r, s, f, s->camera, n, &job)) return e;
        if (int e = enqueue_wave(r, {job}, true, &arena)) return e;
        f->samples += n;
        left -= n;
    }
    return CN_OK;
}
