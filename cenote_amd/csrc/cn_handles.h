// cn_handles.h — the opaque handles behind the C ABI and the host-side
// helpers the ABI's translation units share.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cenote_hip.h"
#include "cn_launch.h"
#include "cn_prep.h"
#include "cn_records.h"

namespace cn {

// Sets the thread-local message cn_last_error() returns; returns `status`.
int fail(int status, const char* fmt, ...);

#define CN_HIP(call)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (call);                                                                            \
        if (e_ != hipSuccess) {                                                                            \
            return ::cn::fail(e_ == hipErrorOutOfMemory ? CN_E_ALLOC : CN_E_DEVICE_FAULT, "%s: %s (%s:%d)", \
                              #call, hipGetErrorString(e_), __FILE__, __LINE__);                          \
        }                                                                                                  \
    } while (0)

// A device allocation that frees itself.
template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t count = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), count(o.count) {
        o.ptr = nullptr;
        o.count = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            reset();
            ptr = o.ptr;
            count = o.count;
            o.ptr = nullptr;
            o.count = 0;
        }
        return *this;
    }
    ~DevBuf() { reset(); }
    void reset() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    hipError_t alloc(size_t n) {
        reset();
        count = n;
        if (n == 0) return hipSuccess;
        return hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T));
    }
    void adopt(T* p, size_t n) {  // take ownership of a device allocation
        reset();
        ptr = p;
        count = n;
    }
    hipError_t upload(const T* src, size_t n) {
        hipError_t e = alloc(n);
        if (e != hipSuccess || n == 0) return e;
        return hipMemcpy(ptr, src, n * sizeof(T), hipMemcpyHostToDevice);
    }
};

// Per-kernel timing: a HIP event pair on the launch stream around every
// launch, read back and summed on demand.
struct EventProfiler final : Profiler {
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, size_t>> open;
    std::vector<std::tuple<int, size_t, size_t>> pairs;
    size_t used = 0;
    double total_ms[K_KINDS] = {};
    uint64_t launches[K_KINDS] = {};

    hipEvent_t next();
    void begin(int kind) override;
    void end(int kind) override;
    void collect();
    ~EventProfiler() override;
};

}  // namespace cn

struct cn_context {
    int device = 0;
    uint32_t compute_units = 0;
    std::string name;
    cn::DevBuf<float> tables;  // the BSDF tables, shared by every scene
};

struct cn_scene {
    cn_context* ctx = nullptr;
    uint64_t serial = 0;  // unique per scene created in this process (graph-cache keys)
    cn_camera camera{};
    cn::DevBuf<cn::BvhNode4Q> nodes;        // what the traversal reads
    cn::DevBuf<cn::BvhNode4P> pnodes;       // the same nodes decoded for camera-ray packets
    std::vector<cn::BvhNode4> host_nodes;    // the full-precision tree (cn_scene_bvh)
    cn::DevBuf<cn::BvhNode8Q> nodes8;       // the 8-wide tree over the same leaves (may be empty)
    std::vector<cn::BvhNode8Q> host_nodes8;
    cn::DevBuf<cn::BvhTriangle> tris;
    cn::DevBuf<cn::InstanceRecord> instances;
    cn::DevBuf<cn_material> materials;
    cn::DevBuf<cn_light_record> lights;
    cn::DevBuf<float> positions, normals;
    cn::DevBuf<uint32_t> indices;
    cn::DevBuf<float4> env_texels;
    cn::DevBuf<float> env_marginal, env_conditional, env_pdfs;
    cn::DevBuf<uint32_t> env_marginal_guide, env_conditional_guide;
    std::vector<cn_light_record> host_lights;
    cn::EnvironmentTables host_env;
    uint32_t env_width = 0, env_height = 0;
    // What an incremental update (cn_scene_update) rebuilds from: the light
    // list's inputs and total power, and the placement it checks edits
    // against (per-instance records and mesh, per-mesh vertex and triangle
    // counts).
    std::vector<cn::EmissiveTriangle> host_emitters;
    std::vector<cn_delta_light> host_deltas;
    double light_power = 0.0;
    std::vector<cn::InstanceRecord> host_records;
    std::vector<uint32_t> host_instance_mesh;
    std::vector<uint32_t> host_mesh_shape;
    cn_scene_info info{};
    cn::SceneView view{};
};

struct cn_film {
    cn_context* ctx = nullptr;
    uint32_t width = 0, height = 0;
    uint32_t samples = 0;
    cn::DevBuf<float4> beauty_sum, albedo_sum, normal_sum;
    cn::DevBuf<float> depth_sum;
    cn::DevBuf<float> moment_sum;  // squared-luminance sums, for noise targets (allocated on demand)
    // A one-sample staging set for cn_film_set_sample / cn_film_accumulate_sample.
    cn::DevBuf<float4> beauty, albedo, normal;
    cn::DevBuf<float> depth;
    // Resolve outputs, allocated on first use.
    cn::DevBuf<float4> out_beauty, out_albedo, out_normal;
    cn::DevBuf<float> out_depth;
};

namespace cn {

// One job of a wave as the host schedules it: `samples` consecutive samples
// from `first_sample` of one view.
struct HostJob {
    cn_scene* scene = nullptr;
    CameraView cam{};
    uint32_t width = 0, height = 0;
    uint32_t first_sample = 0, samples = 0;
    uint32_t max_bounces = 0;  // the job's depth cap; 0 = the renderer's
    cn_film* film = nullptr;   // null: trace only, no accumulate
    bool moments = false;      // also fold squared luminance (the film must hold a moment buffer)
};

// Device memory a wave can run in instead of the renderer's own buffers:
// the path pool with its queues and shadow records, and the slot targets
// (the traversal spill area and counters stay the renderer's).
struct WaveArena {
    PathPool pool;
    WaveTargets targets;  // radiance, albedo, normal, depth, guide: `slots` entries each
    uint32_t capacity;    // paths per pool range
    size_t slots;         // (view pixel, sample) slots one wave may hold
};

// Enqueue one wave of jobs on the renderer's stream (no host sync unless a
// buffer must grow or a job-table ring entry is still in flight). With an
// arena the wave runs in it and nothing is allocated; a wave of more slots
// than the arena holds fails with CN_E_INVALID. The wave's accumulate
// waits for the `after_count` events of `after` (waves on other streams
// that fold into the same films), so films still see their samples in
// order while the waves' tracing overlaps.
int enqueue_wave(cn_renderer* r, std::vector<HostJob> jobs, bool aov, const WaveArena* arena = nullptr,
                 const hipEvent_t* after = nullptr, uint32_t after_count = 0);

// A film job for `samples` more samples of `camera` on `s`, continuing at
// the film's sample count.
int film_job(cn_renderer* r, cn_scene* s, cn_film* f, const cn_camera& camera, uint32_t samples, HostJob* job);

// Samples of a `pixels`-pixel view that fill one wave of the renderer's pool.
uint32_t samples_per_wave(const cn_renderer* r, uint32_t pixels, uint32_t wanted);

// Takes the device address of the allocation-free accumulate's arena on
// `device` (which must be current); once per device, from cn_context_create.
int static_arena_init(int device);

}  // namespace cn

// Job tables in flight: a ring of device tables, each filled from pinned
// host staging and released by an event once its wave has finished.
constexpr uint32_t JOB_RING = 32;

// One captured wave: its launch sequence as an executable graph, keyed by
// everything that reaches a kernel argument.
struct WaveGraph {
    std::vector<uint64_t> key;
    hipGraphExec_t exec = nullptr;
    uint64_t last_use = 0;
};

struct cn_renderer {
    cn_context* ctx = nullptr;
    std::unique_ptr<cn::EventProfiler> profiler;
    uint32_t capacity = 0;
    uint32_t max_bounces = 0;
    uint32_t light_sampling = 0;
    hipStream_t stream = nullptr;
    cn::LaunchShape shape{};
    int counting = 0;
    // path pool
    cn::DevBuf<float4> origin, direction, throughput;
    cn::DevBuf<uint4> hit;
    cn::DevBuf<uint2> key;
    cn::DevBuf<uint32_t> state, ray_queue0, ray_queue1, hit_queue, miss_queue;
    cn::DevBuf<float4> shadow_origin, shadow_dir, shadow_contrib;
    cn::DevBuf<uint32_t> shadow_prim, spill;
    cn::DevBuf<unsigned long long> traversal;
    cn::DevBuf<uint32_t> counters;
    // slot-indexed wave targets
    cn::DevBuf<float4> radiance, albedo, normal, guide;
    cn::DevBuf<float> depth;
    // job-table ring
    cn::DevBuf<cn::JobRecord> jobs;
    cn::JobRecord* staging = nullptr;  // pinned, JOB_RING * MAX_JOBS
    hipEvent_t ring_event[JOB_RING] = {};
    bool ring_used[JOB_RING] = {};
    uint32_t ring_next = 0;
    uint32_t last_ranges = 0;
    cn::PathPool pool{};
    // hipGraph per wave shape: waves replay a captured launch sequence when
    // their shape repeats (progressive frames of the same views).
    bool graphs = false;
    std::vector<WaveGraph> graph_cache;
    uint64_t graph_clock = 0;
    uint64_t graph_hits = 0, graph_captures = 0;
    void drop_graphs() {
        for (WaveGraph& g : graph_cache)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        graph_cache.clear();
    }
    ~cn_renderer() { drop_graphs(); }
};
