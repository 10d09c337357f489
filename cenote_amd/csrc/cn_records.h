// cn_records.h — the records the stage kernels read from HBM, shared by the
// host (which builds them) and the device (which reads them).
//
// Layouts are chosen for gfx950: every record is a whole number of 16-byte
// lanes so a wave64 fetches it with dwordx4 loads, and the per-path state is
// structure-of-arrays so consecutive lanes touch consecutive addresses.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

#include "../../include/cenote_hip.h"

namespace cn {

// The stage queues of one (range, bounce).
enum Queue : uint32_t { Q_RAY = 0, Q_HIT = 1, Q_MISS = 2, Q_SHADOW = 3, Q_COUNT = 4 };

// Each stage queue is split into QUEUE_SPLIT sub-queues, each with its own
// u32 counter alone on a 128-byte line. Returning atomics on one line
// serialize (about 88 M/s on MI355X, against 700 M/s over 8 lines), and with
// one counter per queue the first bounce's claims, not its traversal, set
// the trace kernel's time. A producer wave claims in the sub-queue of its
// 64-item input chunk (chunk index mod QUEUE_SPLIT); sub-queue s keeps its
// entries at [s * sub_capacity, s * sub_capacity + count_s) of the queue's
// array, and a consumer walks the sub-queues one after another. Measured on
// the chart and 2M-triangle scenes: 4 and 8 sub-queues alike, 16 to 64
// slower (each consumer item walks the counts); 8 keeps twice the claim
// headroom of 4.
constexpr uint32_t QUEUE_SPLIT = 8;
constexpr uint32_t QUEUE_CHUNK = 64;      // input items per claiming wave
constexpr uint32_t COUNTER_STRIDE = 32;   // u32 words per counter: one 128-byte line

// Most pixel ranges a single wave may walk (pixels / capacity, rounded up).
constexpr uint32_t MAX_RANGES = 64;
// Counters one range holds: Q_COUNT x QUEUE_SPLIT per bounce, plus one
// spare bounce, then QUEUE_SPLIT striped item counters for the tail launch.
constexpr uint32_t counter_lines_per_range(uint32_t max_bounces) {
    return (max_bounces + 1) * Q_COUNT * QUEUE_SPLIT + QUEUE_SPLIT;
}
// The first of the tail launch's striped item counters (the others follow,
// COUNTER_STRIDE words apart).
constexpr uint32_t tail_claim_index(uint32_t max_bounces) {
    return (max_bounces + 1) * Q_COUNT * QUEUE_SPLIT * COUNTER_STRIDE;
}
constexpr uint32_t counters_per_range(uint32_t max_bounces) {
    return counter_lines_per_range(max_bounces) * COUNTER_STRIDE;
}
__host__ __device__ constexpr uint32_t counter_index(uint32_t bounce, uint32_t queue, uint32_t split) {
    return ((bounce * Q_COUNT + queue) * QUEUE_SPLIT + split) * COUNTER_STRIDE;
}
// Entries one sub-queue can receive from an input of at most `items`
// entries: the chunks k < ceil(items / 64) with k = s (mod QUEUE_SPLIT), at
// most 64 outputs each.
__host__ __device__ constexpr uint32_t sub_capacity_for(uint32_t items) {
    return QUEUE_CHUNK * ((items + QUEUE_CHUNK * QUEUE_SPLIT - 1) / (QUEUE_CHUNK * QUEUE_SPLIT));
}

// Floats in the BSDF table blob (crates/cenote/src/tables.rs:67-80).
constexpr uint32_t BSDF_TABLE_FLOATS = 16928;

constexpr uint32_t LIGHT_NONE = 0xffffffffu;    // shaders/lights.slang:20
constexpr uint32_t MEDIUM_NONE = 0xffffffu;     // shaders/pathstate.slang:188
constexpr uint32_t MASK_CAMERA = 0x01u;         // scene/mod.rs:54-57
constexpr uint32_t MASK_ALL = 0xffu;
constexpr uint32_t LIGHT_KIND_TRIANGLE = 0;     // shaders/lights.slang:23-25
constexpr uint32_t LIGHT_KIND_DISTANT = 1;
constexpr uint32_t LIGHT_KIND_POINT = 2;

// ACEScg luminance weights (color.rs:29-31).
constexpr float LUMA_R = 0.2722287f, LUMA_G = 0.6740818f, LUMA_B = 0.0536895f;

// --- BVH -------------------------------------------------------------------
//
// A 4-wide BVH over world-space triangles (the flattened TLAS x BLAS of
// gpu/accel.rs:82-163). 128 bytes per node: one L2 line on gfx950. The six
// float4s hold the four children's slab bounds as structure-of-arrays, so
// the four box tests are plain lane-local vector math. `child` encodes:
//   bit 31 clear        -> internal node index
//   bit 31 set          -> leaf: bits 0..26 first triangle, bits 27..30
//                          count-1 (1..16 triangles)
//   0xffffffff          -> empty slot (its box is inverted, never hit)
struct alignas(16) BvhNode4 {
    float4 lo_x, hi_x, lo_y, hi_y, lo_z, hi_z;
    uint4 child;
    uint4 meta;  // x = node depth; y..w unused
};
static_assert(sizeof(BvhNode4) == 128, "one 128-byte line per node");

// The node the traversal reads: the same four children with their bounds
// quantized to 8 bits per plane on a per-node grid, 64 bytes (four dwordx4
// fetches, half the full-precision node). Child c's box on axis k spans
//   origin_k + q_lo * step_k  ..  origin_k + q_hi * step_k
// with q the byte c of the plane's word and step_k a power of two (a normal
// float, stored as one so the traversal scales by it with one multiply);
// the host rounds every plane outward (floor / ceil in exact arithmetic),
// so each quantized box contains the full-precision one.
struct alignas(16) BvhNode4Q {
    float ox, oy, oz, step_x;
    uint32_t lo_x, hi_x, lo_y, hi_y;  // byte c: child c's plane, in grid steps from the origin
    uint32_t lo_z, hi_z;
    float step_y, step_z;
    uint4 child;
};
static_assert(sizeof(BvhNode4Q) == 64, "half a 128-byte line per node");

// The same node as a wave of camera rays walks it (128 bytes, decoded from
// the quantized node on the device): child c's plane on axis k at
// origin_k + off_k[c], off = q * step exact in f32 (an 8-bit q times a
// power of two), so a lane meets a plane with one fma. The wave fetches a
// node once, through the scalar unit, for all its lanes, and picks the near
// and far plane words there when its rays' direction signs agree. An empty
// slot's box spans +inf .. -inf on every axis.
struct alignas(16) BvhNode4P {
    float ox, oy, oz, pad;
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    uint32_t child[4];
};
static_assert(sizeof(BvhNode4P) == 128, "one 128-byte line per packet node");
// The packet walk addresses the plane arrays by word offset.
static_assert(offsetof(BvhNode4P, lo_x) == 4 * 4 && offsetof(BvhNode4P, hi_x) == 8 * 4 &&
                  offsetof(BvhNode4P, lo_y) == 12 * 4 && offsetof(BvhNode4P, hi_y) == 16 * 4 &&
                  offsetof(BvhNode4P, lo_z) == 20 * 4 && offsetof(BvhNode4P, hi_z) == 24 * 4 &&
                  offsetof(BvhNode4P, child) == 28 * 4,
              "packet node word layout");

// The 8-wide traversal node: eight children on the same kind of per-node
// grid, one 128-byte line. Byte c of each plane pair (lo_x[0] holds
// children 0..3, lo_x[1] children 4..7) is child c's plane in grid steps;
// the traversal reads the first 96 bytes (six dwordx4 fetches). Collapsed
// from the same binary tree as the 4-wide nodes, over the same leaves.
struct alignas(16) BvhNode8Q {
    float ox, oy, oz;
    uint32_t exps;
    uint32_t lo_x[2], hi_x[2], lo_y[2], hi_y[2], lo_z[2], hi_z[2];
    uint32_t child[8];
    uint32_t depth, pad[7];
};
static_assert(sizeof(BvhNode8Q) == 128, "one 128-byte line per 8-wide node");

constexpr uint32_t CHILD_EMPTY = 0xffffffffu;
constexpr uint32_t CHILD_LEAF = 0x80000000u;
constexpr uint32_t LEAF_MAX = 16;
constexpr uint32_t LEAF_FIRST_MASK = (1u << 27) - 1;

__host__ __device__ inline uint32_t leaf_encode(uint32_t first, uint32_t count) {
    return CHILD_LEAF | ((count - 1) << 27) | first;
}

// One triangle in leaf order, world space, 48 bytes:
//   v0.w = instance (bits), v1.w = primitive (bits),
//   v2.w = flags (bits): bit 0 set = non-opaque instance; bits 8..15 = mask;
//          bits 16..20 = the material class (which lobes the closure has).
struct alignas(16) BvhTriangle {
    float4 v0, v1, v2;
};
static_assert(sizeof(BvhTriangle) == 48, "three dwordx4 fetches per triangle");

constexpr uint32_t TRI_NON_OPAQUE = 1u;
constexpr uint32_t TRI_CLASS_SHIFT = 16;  // bits 16..20: the instance material's class

// A hit-queue entry carries the path index in its low HIT_PATH_BITS bits and
// the hit's material class above them, so shading can order hits by class
// without a dependent load. Pools hold at most CN_POOL_CAPACITY_MAX paths.
constexpr uint32_t HIT_PATH_BITS = 27;
constexpr uint32_t HIT_PATH_MASK = (1u << HIT_PATH_BITS) - 1;
static_assert(CN_POOL_CAPACITY_MAX <= (1u << HIT_PATH_BITS), "path indices fit a hit-queue entry");

// --- Scene records -----------------------------------------------------------

// Per instance, 128 bytes: the GeometryRecord of scene/mod.rs:104-123 with
// the mesh pointers replaced by offsets into the scene's pooled vertex and
// index arrays.
struct alignas(16) InstanceRecord {
    float4 o2w[3];   // rows of the object-to-world transform
    float4 w2o[3];   // rows of its inverse
    uint32_t light;  // first light record, or LIGHT_NONE
    uint32_t vertex_base;
    uint32_t index_base;  // in triangles
    uint32_t flags;       // bits 0..7 = visibility mask, bit 8 = opaque
    uint32_t material_class;     // which lobes the closure has (material-sorted shading)
    int32_t log2_transmittance;  // log2(1 - opacity) in 2^-24 units (log2_transmittance_fixed)
    uint32_t pad[2];
};
static_assert(sizeof(InstanceRecord) == 128, "two dwordx4 lines per instance");

static_assert(sizeof(cn_material) == 144, "material.rs:116");
static_assert(sizeof(cn_light_record) == 80, "lights.rs:115");

// Everything the stage kernels read of a scene, passed by value.
struct SceneView {
    const BvhNode4Q* nodes;
    const BvhNode8Q* nodes8;  // the 8-wide tree over the same leaves (null when the scene has none)
    const BvhNode4P* pnodes;  // `nodes` decoded for wave-packet walks, node for node
    const BvhTriangle* tris;
    const InstanceRecord* instances;
    const cn_material* materials;
    const cn_light_record* lights;
    const float* positions;  // pooled object-space xyz
    const float* normals;    // pooled object-space xyz
    const uint32_t* indices; // pooled, mesh-local vertex indices
    const float* tables;     // BSDF tables (16,928 floats)
    const float4* env_texels;
    const float* env_marginal;
    const float* env_conditional;
    const float* env_pdfs;
    const uint32_t* env_marginal_guide;     // env_marginal_buckets + 1 entries
    const uint32_t* env_conditional_guide;  // env_conditional_buckets + 1 per row
    uint32_t light_count;
    uint32_t env_width;
    uint32_t env_height;
    float env_select_prob;
    uint32_t env_marginal_buckets;
    uint32_t env_conditional_buckets;
};

// The order of a view's camera paths in the slot space: 8-row bands top to
// bottom, each walked column by column, so a wave's 64 consecutive slots
// cover an 8 x 8 block of pixels (fewer rows in the last band of a height
// that is not a multiple of 8). Compact blocks keep a wave's camera rays,
// and the hits and rays that follow them, close together in the BVH and in
// memory. band_pixel maps a slot-order index to its pixel, band_index back.
__host__ __device__ inline uint32_t band_pixel(uint32_t index, uint32_t width, uint32_t height) {
    const uint32_t band = index / (width * 8u);
    const uint32_t j = index - band * (width * 8u);
    const uint32_t left = height - band * 8u;
    const uint32_t rows = left < 8u ? left : 8u;
    const uint32_t x = j / rows;
    return (band * 8u + (j - x * rows)) * width + x;
}
__host__ __device__ inline uint32_t band_index(uint32_t pixel, uint32_t width, uint32_t height) {
    const uint32_t y = pixel / width, x = pixel - y * width;
    const uint32_t band = y / 8u;
    const uint32_t left = height - band * 8u;
    const uint32_t rows = left < 8u ? left : 8u;
    return band * (width * 8u) + x * rows + (y - band * 8u);
}

// A job's slot layout: its pixels (in band order) go in groups of
// SLOT_GROUP consecutive indices, and a group's slots hold its samples one
// after another, sample k of the group's pixel j at k * g + j (g the
// group's pixel count: SLOT_GROUP, fewer only in a job's last group). With
// groups of one pixel each pixel's samples are consecutive, so a wave64 of
// camera paths holds one to a few neighbouring pixels' samples, whose rays
// walk nearly the same nodes; a 1080p frame traced this way took 15.2 ms
// against 16.0 with groups of 8 pixels and 16.9 with the samples of a view
// one after another (every pixel of sample 0, then of sample 1, ...). With
// one sample the layout is band order itself.
#ifndef SLOT_GROUP
#define SLOT_GROUP 1u
#endif
__host__ __device__ inline uint32_t job_slot(uint32_t index, uint32_t k, uint32_t pixels, uint32_t samples) {
    const uint32_t first = index - index % SLOT_GROUP;
    const uint32_t left = pixels - first;
    const uint32_t g = left < SLOT_GROUP ? left : SLOT_GROUP;
    return first * samples + k * g + (index - first);
}
// The inverse: slot `local` of a job -> (band-order index, sample).
__host__ __device__ inline void job_unslot(uint32_t local, uint32_t pixels, uint32_t samples, uint32_t* index,
                                           uint32_t* k) {
    const uint32_t full = pixels / SLOT_GROUP;  // groups of SLOT_GROUP pixels, then one shorter group
    const uint32_t span = SLOT_GROUP * samples;
    const uint32_t group = local / span;
    const uint32_t first = (group < full ? group : full) * SLOT_GROUP;
    const uint32_t g = group < full ? SLOT_GROUP : pixels - first;
    const uint32_t r = local - first * samples;
    *k = r / g;
    *index = first + (r - *k * g);
}

// The SoA path pool (shaders/pathstate.slang:207-232). A wave is a set of
// jobs — (view, consecutive samples) pairs — laid end to end in a slot
// space; a path's target slot is its range base plus its pool index, and
// its sampler key (pixel within its view, sample index) rides in `key`.
struct PathPool {
    float4* origin;       // xyz origin, w = the path's depth cap (bits)
    float4* direction;    // xyz direction, w = cosine to the camera forward
    uint4* hit;           // instance, primitive, b1 bits, b2 bits
    float4* throughput;   // xyz weight, w = previous scatter pdf
    uint32_t* state;      // lobe << 24 | medium
    uint2* key;           // pixel within the view, sample index
    uint32_t* ray_queue[2];  // ping-pong ray queues for bounces >= 1
    uint32_t* hit_queue;
    uint32_t* miss_queue;
    // Shadow records (shaders/pathstate.slang:306-317) as SoA, indexed like
    // a queue's entries (QUEUE_SPLIT sub-queues of sub_capacity records):
    float4* shadow_origin;   // xyz, w = tmax
    float4* shadow_dir;      // xyz, w = pixel (bits)
    float4* shadow_contrib;  // xyz, w = target instance (bits)
    uint32_t* shadow_prim;
    uint32_t* stack_overflow;  // per-lane spill area of the traversal stack
    unsigned long long* traversal_counts;  // when counting: 4 counters per bounce, [bounce][ray nodes,
                                           // ray triangles, shadow nodes, shadow triangles]
    uint32_t sub_capacity;     // entries per sub-queue: sub_capacity_for(pool paths)
};

// The wave's per-slot targets (shaders/pathstate.slang:241-258 plus the
// radiance accumulator): one entry per (view pixel, sample) of the wave.
struct WaveTargets {
    float4* radiance;
    float4* albedo;  // NULL when AOVs are off
    float4* normal;
    float* depth;
    float4* guide;
};

// The camera as raygen reads it (wavefront.rs:84-119 RaygenParams).
struct CameraView {
    float3 position;
    float3 right;
    float3 up;
    float3 forward;
    float aperture;
};

// One job of a wave: `samples` consecutive samples, starting at
// `first_sample`, of every pixel of one view, occupying slots
// [slot_base, slot_base + samples * pixels) as laid out by job_slot. When
// `beauty_sum` is set, the accumulate kernel folds the job's slots into
// that film's sums in sample order (render/film.rs:66-127,
// accumulate.slang:46-67).
constexpr uint32_t JOB_RESET = 1u;  // the job's first sample overwrites the sums

struct JobRecord {
    CameraView cam;
    uint32_t width, height, pixels;
    uint32_t slot_base, pixel_base;
    uint32_t first_sample, samples;
    uint32_t flags;
    uint32_t max_bounces;  // the job's depth cap (paths carry it in origin.w)
    float4* beauty_sum;
    float4* albedo_sum;
    float4* normal_sum;
    float* depth_sum;
    float* moment_sum;  // when set: the running sum of each sample's squared luminance
};

// Jobs one wave may carry.
constexpr uint32_t MAX_JOBS = 256;

// The allocation-free accumulate's device arena: waves of at most
// STATIC_PATHS paths and slots, in one array of STATIC_SCRATCH_BYTES per
// device (path state 76 B, queues and shadow records 68 B, slot targets
// 68 B per path: 53 MiB).
constexpr uint32_t STATIC_PATHS = CN_STATIC_MAX_PIXELS;
constexpr size_t STATIC_SCRATCH_BYTES = size_t(53) << 20;

}  // namespace cn
