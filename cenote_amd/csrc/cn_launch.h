// cn_launch.h — the host-callable launchers of the stage kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "cn_records.h"

namespace cn {

// Traversal stack: STACK_LDS entries per lane in LDS, the rest (up to
// STACK_MAX) in the per-lane HBM spill area. The BVH builder rejects trees
// whose closest-first traversal could need more than STACK_MAX entries
// (3 per level of a 4-wide tree).
constexpr uint32_t STACK_LDS = 16;
constexpr uint32_t STACK_MAX = 128;
constexpr uint32_t STACK_SPILL = STACK_MAX - STACK_LDS;
// Lanes per workgroup of the traversal kernels (trace, first hits).
constexpr uint32_t TRAVERSAL_BLOCK = 256;
// Rows of the traversal counters: one per bounce (depth caps are < 256).
constexpr uint32_t TRAVERSAL_ROWS = 256;

// shade_surface launch variants: workgroup size and where the BSDF tables live.
// The _SORTED variants order each workgroup's hits by material class first.
enum ShadeVariant : uint32_t { SHADE_1024_LDS = 0, SHADE_512_LDS = 1, SHADE_256_LDS = 2, SHADE_256_GLOBAL = 3,
                               SHADE_512_LDS_SORTED = 4, SHADE_256_LDS_SORTED = 5, SHADE_256_GLOBAL_SORTED = 6,
                               SHADE_VARIANTS = 7 };

inline uint32_t shade_block(uint32_t variant) {
    return variant == SHADE_1024_LDS ? 1024u : (variant == SHADE_512_LDS || variant == SHADE_512_LDS_SORTED) ? 512u : 256u;
}

// Grid sizes: the traversal kernels run a fixed, chip-filling grid and
// walk their queue; the shading kernels are capped likewise.
struct LaunchShape {
    uint32_t trace_blocks;   // blocks of 256 lanes for trace
    uint32_t shade_variant;  // a ShadeVariant
    uint32_t shade_blocks;   // most blocks (of the variant's size) for shade_surface
    uint32_t light_blocks;   // blocks of 256 lanes for shade_miss
    uint32_t bvh_width;      // 4 or 8: the tree the traversal kernels walk
    uint32_t tail_bounce;    // 0: every bounce runs as wavefront stages; b >= 1: after bounce b's
                             // trace, one tail launch carries each live path to its end
};

// The film's buffers (render/film.rs:66-127): one wave's sample
// accumulators and the running sums.
struct FilmView {
    float4* beauty;
    float4* beauty_sum;
    float4* albedo;
    float4* albedo_sum;
    float4* normal;
    float4* normal_sum;
    float* depth;
    float* depth_sum;
};

// The stage kernels, for per-kernel timing.
// K_TRACE is intersect fused with the previous bounce's trace_shadow;
// K_TAIL carries the paths alive after the tail bounce's trace to their end.
enum KernelKind { K_RAYGEN = 0, K_TRACE, K_SHADE_MISS, K_SHADE_SURFACE, K_ACCUMULATE, K_TAIL, K_KINDS };

// Brackets each launch (the renderer records HIP events on the launch
// stream when profiling is on).
struct Profiler {
    virtual void begin(int kind) = 0;
    virtual void end(int kind) = 0;
    virtual ~Profiler() = default;
};

// Starts a range of a wave: resets its queue counters.
hipError_t launch_raygen(hipStream_t stream, uint32_t* counters, uint32_t counter_lines, Profiler* prof);

// One bounce: trace, then shade_miss and shade_surface — or, with `tail`,
// trace and then the tail launch that finishes every live path (its item
// counters at `tail_claims`, zeroed with the range's other counters; its
// grid the waves `compute_units` CUs hold at once). Bounce 0's trace
// generates the camera paths of slots [base, base + range_count) of a wave
// whose jobs (sorted by slot base) are `jobs[0 .. job_count)`.
hipError_t launch_bounce(hipStream_t stream, const LaunchShape& shape, const SceneView& s, const PathPool& p,
                         const WaveTargets& tg, uint32_t* counters, const JobRecord* jobs, uint32_t job_count,
                         uint32_t base, uint32_t range_count, uint32_t bounce, uint32_t light_sampling, int counting,
                         bool tail, uint32_t* tail_claims, uint32_t compute_units, Profiler* prof);

// Fold every job's slots into its film sums (jobs sorted by pixel base;
// total_pixels is the sum of the jobs' view pixels).
hipError_t launch_accumulate(hipStream_t stream, const JobRecord* jobs, uint32_t job_count, const WaveTargets& tg,
                             uint32_t total_pixels, Profiler* prof);

// Divides the sums into the FilmView's sample-buffer slots (the caller
// passes the resolve targets there).
hipError_t launch_resolve(hipStream_t stream, const FilmView& f, uint32_t pixels, float samples);

// Per-block partial sums of the estimated relMSE of a film's mean
// (`blocks` doubles; samples >= 2).
hipError_t launch_noise(hipStream_t stream, const float4* sums, const float* moments, uint32_t pixels, float samples,
                        double* partials, uint32_t blocks);

hipError_t launch_sample_sequence(hipStream_t stream, uint32_t pixel, uint32_t dimension, uint32_t count,
                                  float* points, float* values);

hipError_t launch_first_hits(hipStream_t stream, const LaunchShape& shape, const SceneView& s, const CameraView& cam,
                             const PathPool& p, uint32_t width, uint32_t height, uint32_t sample, uint32_t* inst,
                             uint32_t* prim, float* bary);
hipError_t launch_env_sample(hipStream_t stream, const SceneView& s, const float2* u, uint32_t n, float* out);

}  // namespace cn
