// cn_stages.hip — the wavefront stage kernels for gfx950.
//
// One wave traces one sample of every pixel of a target
// (crates/cenote/src/wavefront.rs:1-28): raygen, then per bounce intersect ->
// shade_miss | shade_surface -> trace_shadow, then the film's accumulate.
// Stages talk through device queues whose counts never cross back to the
// host: every kernel after raygen is launched with a fixed, chip-filling
// grid and walks its queue with a grid-stride loop, reading the count the
// previous kernel left (HIP's stand-in for the reference's indirect
// dispatch, wavefront.rs:799-828). Queue pushes are aggregated per wave64 —
// one ballot, one popcount, one atomic per wave, on one of QUEUE_SPLIT
// counters per queue — instead of the one global atomic per push of
// shaders/pathstate.slang:277-283.
//
// Radiance and AOV writes are pixel-owned read-modify-writes: a wave holds
// one path per pixel, and stream order serialises the stages, so the adds
// land in the same per-pixel order on every run (pathstate.slang:165-171).

#include <hip/hip_runtime.h>

#include "cn_launch.h"
#include "cn_math.h"
#include "cn_records.h"
#include "cn_shading.h"
#include "cn_traverse.h"

namespace cn {

constexpr uint32_t BLOCK = 256;
constexpr uint32_t ROULETTE_START = 3;   // shade_surface.slang:53
constexpr uint32_t MATERIAL_CLASSES = 32;  // material_class() values: 5 bits
// A camera path's throughput: unit weight, previous pdf 0, no scatter yet
// (raygen.slang:110).
__device__ __forceinline__ float4 camera_throughput() { return make_float4(1.0f, 1.0f, 1.0f, 0.0f); }

// --- wave64 queue aggregation ------------------------------------------------

// Claim a slot in `counter` for every lane whose `active` is set, with one
// atomic per wave. Works in divergent code: the ballot sees only the lanes
// executing the call, and the leader is one of them.
__device__ __forceinline__ uint32_t wave_claim(uint32_t* counter, bool active) {
    const uint64_t mask = __ballot(active);
    if (mask == 0) return 0;
    const uint32_t lane = __lane_id();
    const uint32_t leader = __ffsll((unsigned long long)mask) - 1;
    const uint32_t below = __popcll(mask & ((1ull << lane) - 1ull));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + below;
}

// The sub-queue a producer lane claims in: that of its wave's 64-item input
// chunk. `first` is the block's first input item (a multiple of 64).
__device__ __forceinline__ uint32_t chunk_split(uint32_t first) {
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / QUEUE_CHUNK);  // wave-uniform
    return (first / QUEUE_CHUNK + wave) % QUEUE_SPLIT;
}

// A consumer sees a queue as its sub-queues laid end to end. The counts are
// reread (scalar loads, cached) wherever they are needed rather than kept
// live through a kernel's loop, which cost registers.
__device__ __forceinline__ uint32_t queue_total(const uint32_t* counters, uint32_t bounce, Queue queue) {
    uint32_t total = 0;
#pragma unroll
    for (uint32_t s = 0; s < QUEUE_SPLIT; ++s) total += counters[counter_index(bounce, queue, s)];
    return total;
}

// Array position of the queue's i-th entry (i < its total).
__device__ __forceinline__ uint32_t queue_slot(const uint32_t* counters, uint32_t bounce, Queue queue, uint32_t i,
                                               uint32_t sub_capacity) {
    uint32_t at = i, base = 0, start = 0;
#pragma unroll
    for (uint32_t s = 0; s + 1 < QUEUE_SPLIT; ++s) {
        start += counters[counter_index(bounce, queue, s)];
        const bool past = i >= start;
        at = past ? i - start : at;
        base = past ? (s + 1) * sub_capacity : base;
    }
    return base + at;
}

__device__ __forceinline__ void wave_add_u64(unsigned long long* target, uint32_t value) {
    uint32_t v = value;
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset);
    if (__lane_id() == 0 && v != 0) atomicAdd(target, (unsigned long long)v);
}

__device__ __forceinline__ void add_pixel(float4* target, uint32_t pixel, v3 c, float a) {
    float4 v = target[pixel];
    v.x += c.x;
    v.y += c.y;
    v.z += c.z;
    v.w += a;
    target[pixel] = v;
}

// A slot's accumulators are not zero-filled by raygen: the bounce-0 kernel
// that takes the camera path (shade_miss or shade_surface, exactly one of
// them) makes each target's first write a store of what adding to zeros
// would give (0 + c, which also turns a -0 into +0 as the add did), and
// stores zeros where it has nothing to add.
__device__ __forceinline__ void put_pixel(float4* target, uint32_t pixel, v3 c, float a, bool first) {
    if (first) target[pixel] = make_float4(0.0f + c.x, 0.0f + c.y, 0.0f + c.z, 0.0f + a);
    else add_pixel(target, pixel, c, a);
}


// --- raygen (raygen.slang:64-112) --------------------------------------------------

__device__ __forceinline__ void camera_ray(const CameraView& cam, uint32_t width, uint32_t height, uint32_t pixel,
                                           uint32_t sample, v3* origin, v3* direction, float* cos_forward) {
    const v3 right = V3(cam.right.x, cam.right.y, cam.right.z);
    const v3 up = V3(cam.up.x, cam.up.y, cam.up.z);
    const v3 forward = V3(cam.forward.x, cam.forward.y, cam.forward.z);
    const f2 jitter = sample_2d(pixel, sample, DIM_CAMERA_JITTER);
    const float x = float(pixel % width) + jitter.x;
    const float y = float(pixel / width) + jitter.y;
    const float u = x / float(width), v = y / float(height);
    const float ndcx = 2.0f * u - 1.0f, ndcy = 1.0f - 2.0f * v;
    const v3 target = (forward + ndcx * right) + ndcy * up;
    v3 o = V3(cam.position.x, cam.position.y, cam.position.z);
    v3 d = normalize(target);
    if (cam.aperture > 0.0f) {
        // Shirley-Chiu concentric disk (raygen.slang:47-62)
        const f2 l = sample_2d(pixel, sample, DIM_CAMERA_LENS);
        const float ox = 2.0f * l.x - 1.0f, oy = 2.0f * l.y - 1.0f;
        float dx = 0.0f, dy = 0.0f;
        if (!(ox == 0.0f && oy == 0.0f)) {
            float r, theta;
            if (fabsf(ox) > fabsf(oy)) {
                r = ox;
                theta = (PI_K / 4.0f) * (oy / ox);
            } else {
                r = oy;
                theta = (PI_K / 2.0f) - (PI_K / 4.0f) * (ox / oy);
            }
            dx = r * cosf(theta);
            dy = r * sinf(theta);
        }
        dx = cam.aperture * dx;
        dy = cam.aperture * dy;
        const v3 offset = dx * normalize(right) + dy * normalize(up);
        o = o + offset;
        d = normalize(target - offset);
    }
    *origin = o;
    *direction = d;
    *cos_forward = dot(d, normalize(forward));
}

// The camera path of wave slot `slot`: its job (jobs sorted by slot base),
// pixel and sample, and the camera ray (raygen.slang:64-112).
struct CameraPath {
    v3 o, d;
    float cos_forward;
    uint2 key;             // (pixel, sample)
    uint32_t max_bounces;  // the job's depth cap
};

__device__ __forceinline__ CameraPath camera_path(const JobRecord* __restrict__ jobs, uint32_t job_count,
                                                  uint32_t slot) {
    uint32_t lo = 0, hi = job_count;
    while (lo + 1 < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (jobs[mid].slot_base <= slot) lo = mid;
        else hi = mid;
    }
    const JobRecord& job = jobs[lo];
    uint32_t index, k;
    job_unslot(slot - job.slot_base, job.pixels, job.samples, &index, &k);
    CameraPath c;
    c.key = make_uint2(band_pixel(index, job.width, job.height), job.first_sample + k);
    camera_ray(job.cam, job.width, job.height, c.key.x, c.key.y, &c.o, &c.d, &c.cos_forward);
    c.max_bounces = job.max_bounces;
    return c;
}

// Starts a range: resets its queue counters (one word per 128-byte counter
// line). The camera paths themselves are generated where bounce 0 traces
// them (k_trace), which stores the records of the paths that hit; the
// reference's raygen pass (raygen.slang:64-112, with Pass::Fill's zero-fill
// of the slot's accumulators, wavefront.rs:733-759) has no separate pass
// here: the bounce-0 kernels store each slot's first values (put_pixel).
__global__ __launch_bounds__(BLOCK) void k_raygen(uint32_t* counters, uint32_t counter_lines) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < counter_lines) counters[i * COUNTER_STRIDE] = 0;
}

// --- shade_miss (shade_miss.slang:27-75) ----------------------------------------------

// One path whose ray at `bounce` escaped: the environment's radiance, MIS
// weighted against next-event estimation, lands in the path's slot, and a
// camera path's AOVs take the miss values.
__device__ __forceinline__ void miss_shade(const SceneView& s, const WaveTargets& tg, uint32_t pixel, v3 direction,
                                           float4 thr, uint32_t bounce, uint32_t light_sampling) {
    float weight = 1.0f;
    if (thr.w > 0.0f) {
        if (light_sampling == CN_LIGHT_SAMPLING_NEE) {
            weight = 0.0f;
        } else if (light_sampling == CN_LIGHT_SAMPLING_MIS) {
            const float neePdf = s.env_select_prob * env_pdf(s, direction);
            weight = power_heuristic(thr.w, neePdf);
        }
    }
    if (tg.albedo) {
        const bool cameraRay = thr.w == 0.0f;
        if (bounce == 0) {
            // A camera path (thr.w == 0): the slot's first AOV writes.
            tg.depth[pixel] = 0.0f + __uint_as_float(0x7f800000u);
            put_pixel(tg.albedo, pixel, S3(1.0f), 0.0f, true);
            tg.normal[pixel] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        } else {
            if (cameraRay) tg.depth[pixel] += __uint_as_float(0x7f800000u);
            const v3 guide = cameraRay ? S3(1.0f) : xyz(tg.guide[pixel]);
            // A path past its first diffuse vertex carries a zero guide.
            // Adding zeros leaves the sum's bits as they are (it starts
            // at +0 and never turns -0), so skip that read-modify-write;
            // any other guide, NaN included, is added as before.
            if (!(guide.x == 0.0f && guide.y == 0.0f && guide.z == 0.0f)) add_pixel(tg.albedo, pixel, guide, 0.0f);
        }
    }
    const v3 contribution = (xyz(thr) * env_radiance(s, direction)) * weight;
    put_pixel(tg.radiance, pixel, contribution, 1.0f, bounce == 0);
}

__device__ __forceinline__ void miss_path(const SceneView& s, const PathPool& p, const WaveTargets& tg, uint32_t path,
                                          uint32_t base, uint32_t bounce, uint32_t light_sampling) {
    const float4 thr = bounce == 0 ? camera_throughput() : p.throughput[path];
    miss_shade(s, tg, base + path, xyz(p.direction[path]), thr, bounce, light_sampling);
}

__global__ __launch_bounds__(BLOCK) void k_shade_miss(SceneView s, PathPool p, WaveTargets tg, uint32_t* counters,
                                                      uint32_t base, uint32_t bounce, uint32_t light_sampling) {
    const uint32_t n = queue_total(counters, bounce, Q_MISS);
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        const uint32_t path = p.miss_queue[queue_slot(counters, bounce, Q_MISS, i, p.sub_capacity)];
        miss_path(s, p, tg, path, base, bounce, light_sampling);
    }
}

// --- trace: intersect (intersect.slang:45-88) fused with the previous
// bounce's trace_shadow (trace_shadow.slang:27-72) ------------------------------
//
// Bounce b's extension rays and bounce b-1's shadow rays are independent, so
// one launch walks both: items [0, n_ray) are rays, [n_ray, n_ray + n_shadow)
// are shadow records. The reference's pass order is kept where it shows:
// a slot's shadow add from bounce b-1 lands before bounce b's shading adds.
// Two half-empty launches per bounce become one fuller one, which is what
// the deep, sparse bounces need.

// The 8-wide walk holds eight keys and refs through its sorting network:
// at eight waves per SIMD (64 VGPRs) it spills, so it runs at fewer.
#ifndef TRACE8_WAVES
#define TRACE8_WAVES 7
#endif
// The tail kernel holds shading's registers and a traversal: three waves
// per SIMD, as shade_surface.
#ifndef TAIL_WAVES
#define TAIL_WAVES 3
#endif

// COUNT: accumulate node and triangle fetches (the counting pass); the
// timed form keeps no counters. Eight waves per SIMD (64 VGPRs; LDS 8 x 16
// KB per CU) fit without scratch: traversal waits on memory, so every
// extra wave hides latency.
// CAMERA: the bounce-0 form, which also shades the camera rays that escape
// (a separate instantiation, so the later bounces' kernel keeps its
// registers).
template <bool COUNT, bool W8, bool CAMERA>
__global__ __launch_bounds__(TRAVERSAL_BLOCK) __attribute__((amdgpu_waves_per_eu(W8 ? TRACE8_WAVES : 8))) void k_trace(SceneView s, PathPool p, WaveTargets tg, uint32_t* counters,
                                                           const JobRecord* __restrict__ jobs, uint32_t job_count,
                                                           uint32_t base, uint32_t range_count, uint32_t bounce,
                                                           uint32_t light_sampling) {
    __shared__ uint32_t lds_stack[STACK_LDS * TRAVERSAL_BLOCK];
    // Bounce 0 traces the camera paths: item i is path i.
    const uint32_t n_ray = CAMERA ? range_count : queue_total(counters, bounce, Q_RAY);
    const uint32_t n = n_ray + (CAMERA ? 0u : queue_total(counters, bounce - 1, Q_SHADOW));
    const uint32_t* rays = p.ray_queue[bounce & 1];
    const uint32_t lanes = gridDim.x * TRAVERSAL_BLOCK;
    Stack st = make_stack(lds_stack, p.stack_overflow, lanes, blockIdx.x * TRAVERSAL_BLOCK + threadIdx.x);
    const uint32_t mask = CAMERA ? MASK_CAMERA : MASK_ALL;
    uint32_t ray_nodes = 0, ray_tris = 0, shadow_nodes = 0, shadow_tris = 0;
    for (uint32_t first = blockIdx.x * TRAVERSAL_BLOCK; first < n; first += lanes) {
        const uint32_t i = first + threadIdx.x;
        const bool is_ray = i < n_ray;
        uint32_t path = 0, cls = 0;
        bool hit = false;
        if (is_ray) {
            v3 o, d;
            uint2 key;
            CameraPath cam;
            if constexpr (CAMERA) {
                // bounce 0 generates the camera path of slot base + i
                path = i;
                cam = camera_path(jobs, job_count, base + i);
                o = cam.o;
                d = cam.d;
                key = cam.key;
            } else {
                path = rays[queue_slot(counters, bounce, Q_RAY, i, p.sub_capacity)];
                o = xyz(p.origin[path]);
                d = xyz(p.direction[path]);
                key = p.key[path];
            }
            TraceHit h;
            if constexpr (CAMERA && !W8 && !COUNT) {
                // Camera rays walk the 4-wide tree as a wave (trace_packet),
                // on the wave's own rows of the stack array. The counting
                // form keeps the per-ray walk, whose fetches the byte model
                // counts.
                static_assert(TRAVERSAL_BLOCK / 64 * PACKET_STACK_ROW <= STACK_LDS * TRAVERSAL_BLOCK,
                              "the packet stack rows fit the block's LDS stack");
                lds_u32* wstack = (lds_u32*)lds_stack + (threadIdx.x / 64) * PACKET_STACK_ROW;
                h = trace_packet(s, o, d, mask, key.x, key.y, wstack, ray_nodes, ray_tris);
            } else {
                h = trace_closest<W8>(s, o, d, 1e30f, mask, st, key.x, key.y, bounce, ray_nodes, ray_tris);
            }
            hit = h.hit;
            if (hit) {
                p.hit[path] = make_uint4(h.inst, h.prim, __float_as_uint(h.b1), __float_as_uint(h.b2));
                if constexpr (CAMERA) {
                    // the path's records, for its shading and later bounces
                    p.origin[path] = F4(o, __uint_as_float(cam.max_bounces));  // w: the path's depth cap
                    p.direction[path] = F4(d, cam.cos_forward);
                    p.key[path] = key;
                }
            } else if constexpr (CAMERA) {
                // A camera ray that escapes is shaded here: no other item of
                // this launch adds to its slot (bounce 0 carries no
                // connections). A camera path's throughput, (1, 1, 1) with no
                // previous pdf, is never stored.
                miss_shade(s, tg, base + path, d, camera_throughput(), 0, light_sampling);
            }
            cls = (h.flags >> TRI_CLASS_SHIFT) & (MATERIAL_CLASSES - 1);
        } else if (i < n) {
            const uint32_t j = queue_slot(counters, bounce - 1, Q_SHADOW, i - n_ray, p.sub_capacity);
            const float4 o4 = p.shadow_origin[j];
            const float4 d4 = p.shadow_dir[j];
            const float4 c4 = p.shadow_contrib[j];
            const uint32_t target = __float_as_uint(c4.w);
            // Any solid non-target hit short of the light's own distance
            // (tmax / 1.001 for a triangle target) lies in front of the
            // target, so it occludes without finding the closest hit.
            const float t_early = target == LIGHT_NONE ? 3.0e38f : o4.w * 0.998f;
            const float att = trace_visibility<W8>(s, xyz(o4), xyz(d4), o4.w, target, p.shadow_prim[j], t_early, st,
                                               shadow_nodes, shadow_tris);
            if (att > 0.0f) add_pixel(tg.radiance, __float_as_uint(d4.w), xyz(c4) * att, 0.0f);
        }
        if (first < n_ray) {  // uniform per block: some wave of this block holds rays
            const uint32_t split = chunk_split(first);
            const uint32_t hs = wave_claim(&counters[counter_index(bounce, Q_HIT, split)], is_ray && hit);
            const uint32_t ms = wave_claim(&counters[counter_index(bounce, Q_MISS, split)], is_ray && !hit);
            if (is_ray) {
                const uint32_t at = split * p.sub_capacity;
                if (hit) p.hit_queue[at + hs] = path | (cls << HIT_PATH_BITS);
                else if (!CAMERA) p.miss_queue[at + ms] = path;  // bounce 0's are counted, and shaded above
            }
        }
    }
    if constexpr (COUNT) {
        unsigned long long* c = p.traversal_counts + 4 * bounce;  // this launch's bounce row
        wave_add_u64(&c[0], ray_nodes);
        wave_add_u64(&c[1], ray_tris);
        wave_add_u64(&c[2], shadow_nodes);
        wave_add_u64(&c[3], shadow_tris);
    }
}

// --- shade_surface (shade_surface.slang:233-521) --------------------------------------

struct Spawn {
    v3 position, below, normal, shading;
    bool frontFace;
};

__device__ __forceinline__ v3 load3(const float* a, uint32_t v) { return V3(a[3 * v], a[3 * v + 1], a[3 * v + 2]); }

// shade_surface.slang:96-181: the hit rebuilt from the object-space mesh,
// offset along the geometric normal by a rigorous bound on its own
// reconstruction error (van Antwerpen 2023).
__device__ inline Spawn spawn_point(const SceneView& s, const InstanceRecord& inst, uint32_t prim, float bx, float by,
                                    v3 rayDirection) {
    const uint32_t* idx = s.indices + 3u * (inst.index_base + prim);
    const uint32_t i0 = inst.vertex_base + idx[0], i1 = inst.vertex_base + idx[1], i2 = inst.vertex_base + idx[2];
    const v3 v0 = load3(s.positions, i0), v1 = load3(s.positions, i1), v2 = load3(s.positions, i2);
    const float4 a0 = inst.o2w[0], a1 = inst.o2w[1], a2 = inst.o2w[2];
    const float4 w0 = inst.w2o[0], w1 = inst.w2o[1], w2 = inst.w2o[2];

    const v3 edge1 = v1 - v0, edge2 = v2 - v0;
    const v3 objPosition = v0 + vfma(S3(bx), edge1, by * edge2);
    const v3 objNormal = cross(edge1, edge2);
    const v3 wldPosition = V3(a0.w + fmaf(a0.x, objPosition.x, fmaf(a0.y, objPosition.y, a0.z * objPosition.z)),
                              a1.w + fmaf(a1.x, objPosition.x, fmaf(a1.y, objPosition.y, a1.z * objPosition.z)),
                              a2.w + fmaf(a2.x, objPosition.x, fmaf(a2.y, objPosition.y, a2.z * objPosition.z)));
    const v3 r0 = xyz(w0), r1 = xyz(w1), r2 = xyz(w2);
    v3 wldNormal = (r0 * objNormal.x + r1 * objNormal.y) + r2 * objNormal.z;
    const float wldScale = rsqrt_exact(dot(wldNormal, wldNormal));
    wldNormal = wldNormal * wldScale;

    const float c0 = 5.9604644775390625e-8f;
    const float c1 = 1.788139769587360206060111522674560546875e-7f;
    const float c2 = 1.19209317972490680404007434844970703125e-7f;
    const v3 extent3 = (vabs(edge1) + vabs(edge2)) + vabs(edge1 - edge2);
    const float extent = fmaxf(fmaxf(extent3.x, extent3.y), extent3.z);
    v3 objErr = vfma(S3(c0), vabs(v0), S3(c1 * extent));
    const v3 aobj = vabs(objPosition);
    const v3 wldErr = vfma(S3(c1), V3(dot(vabs(xyz(a0)), aobj), dot(vabs(xyz(a1)), aobj), dot(vabs(xyz(a2)), aobj)),
                           c2 * vabs(V3(a0.w, a1.w, a2.w)));
    const v3 awld = vabs(wldPosition);
    objErr = vfma(S3(c2),
                  V3(dot(vabs(r0), awld) + fabsf(w0.w), dot(vabs(r1), awld) + fabsf(w1.w),
                     dot(vabs(r2), awld) + fabsf(w2.w)),
                  objErr);
    float wldOffset = dot(wldErr, vabs(wldNormal));
    const float objOffset = dot(objErr, vabs(objNormal));
    wldOffset = fmaf(wldScale, objOffset, wldOffset);

    Spawn sp;
    sp.frontFace = dot(wldNormal, rayDirection) < 0.0f;
    if (dot(wldNormal, rayDirection) > 0.0f) wldNormal = -wldNormal;
    const v3 n0 = load3(s.normals, i0), n1 = load3(s.normals, i1), n2 = load3(s.normals, i2);
    const v3 objShading = (n0 + bx * (n1 - n0)) + by * (n2 - n0);
    v3 shading = normalize((r0 * objShading.x + r1 * objShading.y) + r2 * objShading.z);
    if (dot(shading, wldNormal) < 0.0f) shading = -shading;
    sp.position = vfma(S3(wldOffset), wldNormal, wldPosition);
    sp.below = vfma(S3(-wldOffset), wldNormal, wldPosition);
    sp.normal = wldNormal;
    sp.shading = shading;
    return sp;
}

__device__ __forceinline__ InstanceRecord load_instance(const InstanceRecord* recs, uint32_t i) {
    return recs[i];
}

// The connection a shaded hit sends toward a light (trace_shadow.slang's
// record): origin and the connection's length, direction, the path's slot,
// the unshadowed contribution, and the target light's (instance,
// primitive).
struct ShadowRecord {
    v3 origin;
    float tmax;
    v3 direction;
    uint32_t pixel;
    v3 contrib;
    uint32_t instance;
    uint32_t primitive;
};

// shade_surface for one path's hit at `bounce`: emission into the slot,
// the next-event connection, and the scattered ray. Every lane that gets
// as far as the connection calls `send_shadow(push, record)` (push false
// when there is nothing to connect), so a sink may be a wave-wide claim.
// Returns true when the path goes on (its next ray and throughput are in
// the path records); `below` then tells whether that ray leaves from below
// the surface (transmitted).
template <typename SendShadow>
__device__ __forceinline__ bool shade_path(const SceneView& s, const PathPool& p, const WaveTargets& tg,
                                           const float* tables, uint32_t path, uint32_t base, uint32_t bounce,
                                           uint32_t light_sampling, SendShadow&& send_shadow, bool& below) {
    const bool aovOn = tg.albedo != nullptr;
    bool push_ray = false;
    const uint32_t pixel = base + path;  // the path's target slot
    const uint2 key = p.key[path];
    const uint4 hit = p.hit[path];
    const float4 dir4 = p.direction[path];
    const v3 direction = xyz(dir4);
    const float4 origin4 = p.origin[path];
    const v3 origin = xyz(origin4);
    const uint32_t depth_cap = __float_as_uint(origin4.w);
    const InstanceRecord inst = load_instance(s.instances, hit.x);
    const Mat material = load_material(s.materials, hit.x);
    const Spawn spawn = spawn_point(s, inst, hit.y, __uint_as_float(hit.z), __uint_as_float(hit.w), direction);
    float4 thr = bounce == 0 ? camera_throughput() : p.throughput[path];

    uint32_t medium = bounce == 0 ? MEDIUM_NONE : (p.state[path] & 0xffffffu);
    if (medium != MEDIUM_NONE) {
        const Mat interior = load_material(s.materials, medium);
        if (interior.transmissionDepth > 0.0f) {
            const v3 tc = vclamp(interior.transmissionColor, 1e-4f, 1.0f);
            const v3 sigma = (-V3(logf(tc.x), logf(tc.y), logf(tc.z))) / interior.transmissionDepth;
            const v3 e = -sigma * length(origin - spawn.position);
            thr.x *= expf(e.x);
            thr.y *= expf(e.y);
            thr.z *= expf(e.z);
        }
    }
    const v3 T = xyz(thr);

    v3 contribution = S3(0.0f);
    if (inst.light != LIGHT_NONE && spawn.frontFace) {
        float weight = 1.0f;
        if (thr.w > 0.0f) {
            if (light_sampling == CN_LIGHT_SAMPLING_NEE) {
                weight = 0.0f;
            } else if (light_sampling == CN_LIGHT_SAMPLING_MIS) {
                const v3 delta = spawn.position - origin;
                const float cosLight = fabsf(dot(spawn.normal, direction));
                const float neePdf = (1.0f - s.env_select_prob) * s.lights[inst.light + hit.y].pdf *
                                     dot(delta, delta) / cosLight;
                weight = power_heuristic(thr.w, neePdf);
            }
        }
        contribution = (T * material.emission) * weight;
    }
    // The emission lands now rather than when the path's fate is
    // known; a terminal path adds its sample count (and an emission
    // that was not positive) at the end. A slot sees the same sums.
    const bool lit = any_positive(contribution);
    if (bounce == 0) put_pixel(tg.radiance, pixel, lit ? contribution : S3(0.0f), 0.0f, true);
    else if (lit) add_pixel(tg.radiance, pixel, contribution, 0.0f);

    // The next-event candidate (a point on a light, or an
    // environment direction) depends only on the hit point, so it is
    // drawn before the closure is resolved; only the candidate stays
    // live across the resolution.
    const float envProb = s.env_select_prob;
    const bool nee = bounce + 1 < depth_cap && light_sampling != CN_LIGHT_SAMPLING_BSDF &&
                     (s.light_count > 0 || envProb > 0.0f);
    LightConnection cand{S3(0.0f), 1e30f, S3(0.0f), 0.0f, false, LIGHT_NONE, 0u};
    bool cand_env = false;
    if (nee) {
        const float uLight = sample_1d(key.x, key.y, bounce_dim(bounce, SLOT_NEE_LIGHT));
        const f2 uPoint = sample_2d(key.x, key.y, bounce_dim(bounce, SLOT_NEE_POINT));
        if (uLight < envProb) {
            const EnvSample e = env_sample(s, uPoint);
            cand.direction = e.direction;
            cand.radiance = e.radiance;
            cand.pdf = e.pdf;
            cand_env = true;
        } else {
            const float uList = (uLight - envProb) / (1.0f - envProb);
            cand = sample_lights(s, uList, uPoint, spawn.position);
        }
    }

    const v3 shading = spawn.shading;
    const Frame frame = make_frame(shading);
    const v3 wo = to_local(frame, -direction);
    const Closure closure = resolve_closure(tables, material, wo, medium == hit.x);

    v3 guide = S3(0.0f);
    if (aovOn) {
        guide = bounce == 0 ? S3(1.0f) : xyz(tg.guide[pixel]);
        if (bounce == 0) tg.depth[pixel] = 0.0f + length(origin - spawn.position) * dir4.w;
    }
    const v3 aovAlbedo = vclamp(closure.aovAlbedo, 0.0f, 1.0f);
    bool terminal = false;

    if (bounce + 1 >= depth_cap) {
        terminal = true;
    } else {
        // Next-event estimation (shade_surface.slang:363-446), with the
        // candidate drawn above.
        bool push_shadow = false;
        v3 sh_origin = S3(0.0f), sh_contrib = S3(0.0f);
        if (nee) {
            const v3 wiLight = to_local(frame, cand.direction);
            const bool through = dot(cand.direction, spawn.normal) < 0.0f;
            if (cand.pdf > 0.0f && wiLight.z != 0.0f && (wiLight.z < 0.0f) == through) {
                const float neePdf = (cand_env ? envProb : 1.0f - envProb) * cand.pdf;
                const BsdfEval light = eval_with_pdf(closure, wiLight, wo);
                const float weight = !cand.delta && light_sampling == CN_LIGHT_SAMPLING_MIS
                                         ? power_heuristic(neePdf, light.pdf)
                                         : 1.0f;
                const v3 unshadowed =
                    (((T * cand.radiance) * light.f) * fabsf(wiLight.z)) * (weight / neePdf);
                sh_origin = through ? spawn.below : spawn.position;
                if (any_positive(unshadowed)) {
                    push_shadow = true;
                    sh_contrib = unshadowed;
                }
            }
        }
        // The connection goes out before scattering, so its record
        // is dead through the rest of the shading.
        send_shadow(push_shadow, ShadowRecord{sh_origin, cand.tmax, cand.direction, pixel, sh_contrib,
                                              cand.instance, cand.primitive});

        // Scatter (shade_surface.slang:453-489).
        const float uLobe = sample_1d(key.x, key.y, bounce_dim(bounce, SLOT_BSDF_LOBE));
        const f2 u = sample_2d(key.x, key.y, bounce_dim(bounce, SLOT_BSDF_DIR));
        BsdfSample scatter;
        BsdfEval drawn{S3(0.0f), 0.0f};
        if (sample_bsdf_direction(closure, wo, uLobe, u, scatter)) drawn = eval_with_pdf(closure, scatter.wi, wo);
        scatter.pdf = drawn.pdf;
        const v3 wiWorld = to_world(frame, scatter.wi);
        v3 scattered = S3(0.0f);
        if (scatter.pdf > 0.0f && scatter.wi.z != 0.0f &&
            (scatter.wi.z > 0.0f) == (dot(wiWorld, spawn.normal) > 0.0f)) {
            scattered = (T * drawn.f) * (fabsf(scatter.wi.z) / scatter.pdf);
        }
        float survival = fmaxf(scattered.x, fmaxf(scattered.y, scattered.z));
        if (survival <= 0.0f) {
            terminal = true;
        } else {
            if (bounce >= ROULETTE_START) {
                survival = fminf(1.0f, survival);
                if (sample_1d(key.x, key.y, bounce_dim(bounce, SLOT_RR)) >= survival) terminal = true;
                else scattered = scattered / survival;
            }
        }
        if (!terminal) {
            if (aovOn) {
                const float record = aov_record_fraction(closure, scatter.lobe);
                if (record > 0.0f && any_positive(guide)) {
                    const v3 g = guide * record;
                    put_pixel(tg.albedo, pixel, g * aovAlbedo, 0.0f, bounce == 0);
                    put_pixel(tg.normal, pixel, g * shading, 0.0f, bounce == 0);
                } else if (bounce == 0) {
                    put_pixel(tg.albedo, pixel, S3(0.0f), 0.0f, true);
                    put_pixel(tg.normal, pixel, S3(0.0f), 0.0f, true);
                }
                tg.guide[pixel] = F4((guide * (1.0f - record)) * aov_pass_tint(closure, scatter.lobe), 0.0f);
            }
            if (scatter.lobe == LOBE_TRANSMISSION && material.thinWalled == 0) {
                medium = medium == hit.x ? MEDIUM_NONE : hit.x;
            }
            const bool transmitted = scatter.wi.z < 0.0f;
            below = transmitted;
            p.origin[path] = F4(transmitted ? spawn.below : spawn.position, origin4.w);
            p.direction[path] = F4(wiWorld, 0.0f);
            p.throughput[path] = F4(scattered, scatter.pdf);
            p.state[path] = (scatter.lobe << 24) | (medium & 0xffffffu);
            push_ray = true;
        }
    }
    if (terminal) {
        if (aovOn && any_positive(guide)) {
            put_pixel(tg.albedo, pixel, guide * aovAlbedo, 0.0f, bounce == 0);
            put_pixel(tg.normal, pixel, guide * shading, 0.0f, bounce == 0);
        } else if (aovOn && bounce == 0) {
            put_pixel(tg.albedo, pixel, S3(0.0f), 0.0f, true);
            put_pixel(tg.normal, pixel, S3(0.0f), 0.0f, true);
        }
        add_pixel(tg.radiance, pixel, lit ? S3(0.0f) : contribution, 1.0f);
    }
    return push_ray;
}

// Launch-shape variants of shade_surface, picked per renderer (the
// autotuner measures them). The BSDF tables — 67,712 bytes that each
// closure evaluation reads ~20 times — are staged into LDS once per
// workgroup, so wide groups amortise the staging and a CU's LDS holds two
// groups; the last variant reads them through the caches instead and is not
// limited by LDS.
// Material sorting (the SORT variants): each workgroup's hits are ordered by
// their instance's material class (which lobes the closure has) with an LDS
// counting sort before shading, so a wave64 shades one lobe set where it
// can instead of every branch of the closure. The order in which a
// workgroup shades its hits reaches no result: every add lands in the
// path's own slot.
template <uint32_t SB, bool LDS_TABLES, bool SORT>
__global__ __launch_bounds__(SB) __attribute__((amdgpu_waves_per_eu(3))) void k_shade_surface(SceneView s, PathPool p, WaveTargets tg, uint32_t* counters,
                                                      uint32_t base, uint32_t bounce, uint32_t light_sampling) {
    const uint32_t n = queue_total(counters, bounce, Q_HIT);
    if (blockIdx.x * SB >= n) return;  // uniform per block: nothing to shade here
    const float* tables = s.tables;
    if constexpr (LDS_TABLES) {
        __shared__ float lds_tables[T_FLOATS];
        for (uint32_t k = threadIdx.x; k < T_FLOATS / 4; k += SB) {
            reinterpret_cast<float4*>(lds_tables)[k] = reinterpret_cast<const float4*>(s.tables)[k];
        }
        __syncthreads();
        tables = lds_tables;
    }
    uint32_t* next_rays = p.ray_queue[(bounce + 1) & 1];
    // Sorted variants: the lane that held each sorted hit before the sort.
    __shared__ uint32_t sorted_home[SORT ? SB : 1];

    for (uint32_t first = blockIdx.x * SB; first < n; first += gridDim.x * SB) {
        const uint32_t i = first + threadIdx.x;
        bool valid = i < n;
        const uint32_t entry = valid ? p.hit_queue[queue_slot(counters, bounce, Q_HIT, i, p.sub_capacity)] : 0u;
        uint32_t path = entry & HIT_PATH_MASK;
        if constexpr (SORT) {
            // Bucket MATERIAL_CLASSES holds the lanes past the queue's end.
            __shared__ uint32_t bucket[MATERIAL_CLASSES + 1];
            __shared__ uint32_t sorted[SB];
            const uint32_t cls = valid ? entry >> HIT_PATH_BITS : MATERIAL_CLASSES;
            if (threadIdx.x <= MATERIAL_CLASSES) bucket[threadIdx.x] = 0;
            __syncthreads();
            const uint32_t rank = atomicAdd(&bucket[cls], 1u);
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t run = 0;
                for (uint32_t c = 0; c <= MATERIAL_CLASSES; ++c) {
                    const uint32_t k = bucket[c];
                    bucket[c] = run;
                    run += k;
                }
            }
            __syncthreads();
            sorted[bucket[cls] + rank] = valid ? path : 0xffffffffu;
            sorted_home[bucket[cls] + rank] = threadIdx.x;  // read back after shading
            __syncthreads();
            path = sorted[threadIdx.x];
            valid = path != 0xffffffffu;
            if (!valid) path = 0;
        }
        bool push_ray = false, below = false;
        if (valid) {
            push_ray = shade_path(s, p, tg, tables, path, base, bounce, light_sampling,
                                  [&](bool push, const ShadowRecord& r) {
                                      const uint32_t split = chunk_split(first);
                                      const uint32_t ss =
                                          split * p.sub_capacity +
                                          wave_claim(&counters[counter_index(bounce, Q_SHADOW, split)], push);
                                      if (push) {
                                          p.shadow_origin[ss] = F4(r.origin, r.tmax);
                                          p.shadow_dir[ss] = F4(r.direction, __uint_as_float(r.pixel));
                                          p.shadow_contrib[ss] = F4(r.contrib, __uint_as_float(r.instance));
                                          p.shadow_prim[ss] = r.primitive;
                                      }
                                  },
                                  below);
        }
        if constexpr (SORT) {
            // The next bounce's rays go out in the hits' queue order, not in
            // material order: consecutive lanes of the next kernels then
            // read the records of nearby paths. The block's rays are then
            // stably partitioned, those leaving from above the surface first
            // and the transmitted ones after: more of a trace wave's rays
            // are of one kind and walk similar numbers of nodes, while the
            // records stay within the block's own paths. (Grouping by
            // direction octant as well scattered the next misses' slots and
            // measured slower.)
            __shared__ uint32_t pushed[SB];
            __shared__ uint32_t wave_counts[SB / 64][2];
            constexpr uint32_t BELOW = 1u << 31;  // paths use HIT_PATH_BITS < 31 bits
            pushed[sorted_home[threadIdx.x]] = push_ray ? (path | (below ? BELOW : 0u)) : 0xffffffffu;
            __syncthreads();
            const uint32_t v = pushed[threadIdx.x];
            const bool ray = v != 0xffffffffu;
            const bool down = ray && (v & BELOW);
            const bool up = ray && !down;
            const uint64_t mu = __ballot(up), md = __ballot(down);
            const uint32_t wave = threadIdx.x / 64;
            const uint64_t lt = (1ull << __lane_id()) - 1ull;
            if (__lane_id() == 0) {
                wave_counts[wave][0] = __popcll(mu);
                wave_counts[wave][1] = __popcll(md);
            }
            __syncthreads();
            uint32_t before_up = 0, before_down = 0, total_up = 0, total_down = 0;
            for (uint32_t k = 0; k < SB / 64; ++k) {
                const uint32_t cu = wave_counts[k][0], cd = wave_counts[k][1];
                before_up += k < wave ? cu : 0u;
                before_down += k < wave ? cd : 0u;
                total_up += cu;
                total_down += cd;
            }
            const uint32_t pos = up ? before_up + __popcll(mu & lt)
                                    : down ? total_up + before_down + __popcll(md & lt)
                                           : total_up + total_down + (threadIdx.x - (before_up + before_down) -
                                                                       __popcll((mu | md) & lt));
            __syncthreads();  // every lane has read its entry
            pushed[pos] = ray ? (v & ~BELOW) : 0xffffffffu;
            __syncthreads();
            path = pushed[threadIdx.x];
            push_ray = path != 0xffffffffu;
        }
        const uint32_t split = chunk_split(first);
        const uint32_t rs = wave_claim(&counters[counter_index(bounce + 1, Q_RAY, split)], push_ray);
        if (push_ray) next_rays[split * p.sub_capacity + rs] = path;
    }
}

// --- tail: the paths still alive after one bounce's trace, each carried to its end ----
//
// Past the first few bounces a wave holds few live paths, and each bounce's
// three launches last as long as the slowest of them: the deep bounces cost
// the sum, over bounces, of the slowest path's latency. From a chosen bounce
// on, one launch takes every path of that bounce's miss and hit queues
// through all of its remaining bounces — shading, the next-event connection
// traced at once, the next closest hit — with the wavefront kernels' own
// device functions. A slot sees the same adds in the same order as in the
// wavefront (bounce b's shading, then b's connection, then bounce b + 1),
// so the film is bit for bit the same. Misses come first in item order, so
// most waves hold one kind of work.
// The lanes work in pairs: the even lane shades its path and walks the
// path's next closest hit while its odd mate walks the bounce's next-event
// connection, so the two walks of a bounce overlap instead of following
// one another (a deep path's latency is a chain of dependent fetches; on
// the 2M-triangle scene the tail took 1.54 ms per frame this way against
// 2.21 ms with one lane per path). The owner adds the connection's light
// after both walks and before it shades the next hit.
template <bool W8>
__global__ __launch_bounds__(TRAVERSAL_BLOCK) __attribute__((amdgpu_waves_per_eu(TAIL_WAVES))) void k_tail(
    SceneView s, PathPool p, WaveTargets tg, uint32_t* counters, uint32_t* claims, uint32_t base, uint32_t bounce,
    uint32_t light_sampling) {
    __shared__ uint32_t lds_stack[STACK_LDS * TRAVERSAL_BLOCK];
    const uint32_t n_miss = queue_total(counters, bounce, Q_MISS);
    const uint32_t n = n_miss + queue_total(counters, bounce, Q_HIT);
    const uint32_t lanes = gridDim.x * TRAVERSAL_BLOCK;
    Stack st = make_stack(lds_stack, p.stack_overflow, lanes, blockIdx.x * TRAVERSAL_BLOCK + threadIdx.x);
    uint32_t nodes = 0, tris = 0;  // the walks' fetch counts, not kept here
    const bool helper = threadIdx.x & 1u;
    const uint32_t owner = __lane_id() & ~1u, mate = __lane_id() | 1u;
    // Items: the miss-queue entries [0, n_miss), then the hit-queue entries
    // [n_miss, n), one per lane pair. Pairs take items as they free up: a
    // wave claims one item for each of its idle pairs with one atomic on one
    // of QUEUE_SPLIT striped counters (stripe k deals items k, k + QUEUE_SPLIT,
    // ...), moving to a stripe with items left once its own runs out, so no
    // pair waits on the longest path of its wave while items remain. A pair
    // carries its path from bounce to bounce itself, in order.
    uint32_t stripe = (blockIdx.x * (TRAVERSAL_BLOCK / 64) + threadIdx.x / 64) % QUEUE_SPLIT;  // wave-uniform
    bool items_left = true;  // wave-uniform
    bool live = false;  // the pair carries a path (the same in both lanes)
    uint32_t path = 0, b = 0;
    while (true) {
        const uint64_t idle = __ballot(!live && !helper);  // the idle pairs' owners
        if (idle != 0 && items_left) {
            const uint32_t want = __popcll(idle);
            const uint32_t leader = __ffsll((unsigned long long)idle) - 1;
            uint32_t c = 0;
            if (__lane_id() == leader) c = atomicAdd(&claims[stripe * COUNTER_STRIDE], want);
            c = __builtin_amdgcn_readlane(c, leader);
            if (!live) {
                const uint32_t rank = __popcll(idle & ((1ull << owner) - 1));
                const uint32_t i = stripe + QUEUE_SPLIT * (c + rank);
                if (i < n_miss) {
                    const uint32_t entry = p.miss_queue[queue_slot(counters, bounce, Q_MISS, i, p.sub_capacity)];
                    if (!helper) miss_path(s, p, tg, entry, base, bounce, light_sampling);
                } else if (i < n) {
                    path = p.hit_queue[queue_slot(counters, bounce, Q_HIT, i - n_miss, p.sub_capacity)] &
                           HIT_PATH_MASK;
                    b = bounce;
                    live = true;
                }
            }
            if (stripe + QUEUE_SPLIT * (c + want) >= n) {
                // This stripe has no items left: read every stripe's counter
                // at once and move to one that has (counters only grow, so
                // when none seems to, none has).
                const uint32_t k = __lane_id();
                uint32_t taken = 0xffffffffu;
                if (k < QUEUE_SPLIT)
                    taken = __hip_atomic_load(&claims[k * COUNTER_STRIDE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t open = __ballot(k < QUEUE_SPLIT && k + QUEUE_SPLIT * taken < n);
                if (open == 0) items_left = false;
                else stripe = __ffsll((unsigned long long)open) - 1;
            }
        }
        if (__ballot(live) == 0) {
            if (!items_left) break;
            continue;
        }
        if (live) {  // one bounce of each pair's path
                bool connect = false, next = false, below = false;
                ShadowRecord sh{};
                if (!helper)
                    next = shade_path(s, p, tg, s.tables, path, base, b, light_sampling,
                                      [&](bool push, const ShadowRecord& r) {
                                          connect = push;
                                          sh = r;
                                      },
                                      below);
                // The owner hands its decisions and the connection to its mate.
                connect = __shfl(connect ? 1 : 0, owner) != 0;
                next = __shfl(next ? 1 : 0, owner) != 0;
                sh.origin = V3(__shfl(sh.origin.x, owner), __shfl(sh.origin.y, owner), __shfl(sh.origin.z, owner));
                sh.direction =
                    V3(__shfl(sh.direction.x, owner), __shfl(sh.direction.y, owner), __shfl(sh.direction.z, owner));
                sh.tmax = __shfl(sh.tmax, owner);
                sh.instance = __shfl(sh.instance, owner);
                sh.primitive = __shfl(sh.primitive, owner);
                // The owner walks the next ray, the mate the connection.
                v3 o = sh.origin, d = sh.direction;
                float tmax = sh.tmax;
                uint2 key = make_uint2(0u, 0u);
                if (!helper && next) {
                    o = xyz(p.origin[path]);
                    d = xyz(p.direction[path]);
                    key = p.key[path];
                }
                if (!helper) tmax = 1e30f;
                const SlabRay r = make_slab(o, d);
                const WRay w = make_wray(o, d);
                TraceHit best{false, tmax, 0u, 0u, 0.0f, 0.0f, 0u};
                bool occluded = false, translucent = false;
                // As the trace kernel takes a connection: a solid hit short of a
                // triangle light occludes without the closest hit.
                const float t_early = sh.instance == LIGHT_NONE ? 3.0e38f : sh.tmax * 0.998f;
                if (helper ? connect : next) {
                    walk(tree_of<W8>(s), r, best.t, st, nodes, [&](uint32_t ref) {
                        if (helper) {
                            occluded = visibility_leaf(s, w, ref, sh.instance, sh.primitive, t_early, best, translucent,
                                                       tris);
                            return occluded;
                        }
                        closest_leaf(s, TrisPerLane{s.tris}, w, ref, MASK_ALL, key.x, key.y, b + 1, best, tris);
                        return false;
                    });
                }
                float att = 0.0f;
                if (helper && connect && !occluded && !(best.hit && !(best.inst == sh.instance && best.prim == sh.primitive)))
                    att = translucent ? shadow_attenuation<W8>(s, r, w, best.hit ? best.t : sh.tmax, sh.instance,
                                                               sh.primitive, st, nodes)
                                      : 1.0f;
                att = __shfl(att, mate);
                if (!helper && connect && att > 0.0f) add_pixel(tg.radiance, sh.pixel, sh.contrib * att, 0.0f);
            if (!next) {
                live = false;
            } else {
                const bool hit = __shfl(best.hit ? 1 : 0, owner) != 0;
                if (!hit) {
                    if (!helper) miss_path(s, p, tg, path, base, b + 1, light_sampling);
                    live = false;
                } else {
                    if (!helper)
                        p.hit[path] =
                            make_uint4(best.inst, best.prim, __float_as_uint(best.b1), __float_as_uint(best.b2));
                    ++b;
                }
            }
        }
    }
}


// --- film (accumulate.slang:46-67, resolve.slang:31-42) ----------------------------

__device__ __forceinline__ float4 guarded(float4 c) {
    const bool bad = isnan(c.x) || isnan(c.y) || isnan(c.z) || isnan(c.w) || isinf(c.x) || isinf(c.y) ||
                     isinf(c.z) || isinf(c.w);
    return bad ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : c;
}

__device__ __forceinline__ float4 add4(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }

// The film fold of one pixel: its running sums, updated sample by sample
// in sample order. Each sum sees only its own target's samples, so the
// targets may be folded one after another.
struct PixelSums {
    float4 b, a, nn;
    float d, m;
    bool reset, aov, moments;
    __device__ __forceinline__ void radiance(uint32_t k, float4 cb) {
        const bool first = reset && k == 0;
        cb = guarded(cb);
        b = first ? cb : add4(b, cb);
        if (moments) {
            const float y = (cb.x * LUMA_R + cb.y * LUMA_G) + cb.z * LUMA_B;
            m = first ? y * y : m + y * y;
        }
    }
    __device__ __forceinline__ void albedo(uint32_t k, float4 ca) {
        ca = guarded(ca);
        a = (reset && k == 0) ? ca : add4(a, ca);
    }
    __device__ __forceinline__ void normal(uint32_t k, float4 cn) {
        cn = guarded(cn);
        nn = (reset && k == 0) ? cn : add4(nn, cn);
    }
    __device__ __forceinline__ void depth(uint32_t k, float cd) {
        if (isnan(cd)) cd = 0.0f;
        d = (reset && k == 0) ? cd : d + cd;
    }
};

__device__ __forceinline__ PixelSums load_sums(const JobRecord& job, uint32_t px, bool aov) {
    PixelSums s;
    s.reset = job.flags & JOB_RESET;
    s.aov = aov;
    s.moments = job.moment_sum != nullptr;
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    s.b = s.reset ? zero : job.beauty_sum[px];
    s.a = zero;
    s.nn = zero;
    s.d = 0.0f;
    s.m = (s.moments && !s.reset) ? job.moment_sum[px] : 0.0f;
    if (aov && !s.reset) {
        s.a = job.albedo_sum[px];
        s.nn = job.normal_sum[px];
        s.d = job.depth_sum[px];
    }
    return s;
}

__device__ __forceinline__ void store_sums(const JobRecord& job, uint32_t px, const PixelSums& s) {
    job.beauty_sum[px] = s.b;
    if (s.moments) job.moment_sum[px] = s.m;
    if (s.aov) {
        job.albedo_sum[px] = s.a;
        job.normal_sum[px] = s.nn;
        job.depth_sum[px] = s.d;
    }
}

__device__ __forceinline__ uint32_t job_of(const JobRecord* __restrict__ jobs, uint32_t job_count, uint32_t i) {
    uint32_t lo = 0, hi = job_count;
    while (lo + 1 < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (jobs[mid].pixel_base <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Samples staged per pass of the accumulate's LDS path (16 measured faster
// than 8, and 4 much slower, on the 1080p and 720p frames).
#ifndef ACC_CHUNK
#define ACC_CHUNK 16u
#endif

// One thread per view pixel of the wave: fold the job's samples into its
// film's sums in sample order, reading the sums once. A block takes BLOCK
// consecutive band-order pixels. When they all lie in one job's full slot
// groups (every tile of a one-job wave but its last), their slots are one
// contiguous span, and the block stages them through LDS ACC_CHUNK samples
// at a time: coalesced loads of each group's run of slots, then each
// thread folds its own pixel's samples from LDS in sample order. Other
// tiles fold straight from the slots, one thread per pixel.
__global__ __launch_bounds__(BLOCK) void k_accumulate(const JobRecord* __restrict__ jobs, uint32_t job_count,
                                                      WaveTargets tg, uint32_t total_pixels) {
    __shared__ float4 stage[BLOCK * (ACC_CHUNK + 1)];
    for (uint32_t tile = blockIdx.x * BLOCK; tile < total_pixels; tile += gridDim.x * BLOCK) {
        const JobRecord& tj = jobs[job_of(jobs, job_count, tile)];  // uniform per block
        const uint32_t index0 = tile - tj.pixel_base;
        const bool staged = tj.beauty_sum != nullptr && index0 % SLOT_GROUP == 0 &&
                            index0 + BLOCK <= (tj.pixels / SLOT_GROUP) * SLOT_GROUP;
        if (staged) {
            const JobRecord& job = tj;
            const uint32_t index = index0 + threadIdx.x;
            const uint32_t px = band_pixel(index, job.width, job.height);
            const bool aov = tg.albedo != nullptr && job.albedo_sum != nullptr;
            PixelSums sums = load_sums(job, px, aov);
            const uint32_t S = job.samples;
            const uint32_t tile_slot = job.slot_base + index0 * S;  // the tile's groups, end to end
            for (uint32_t k0 = 0; k0 < S; k0 += ACC_CHUNK) {
                const uint32_t c = S - k0 < ACC_CHUNK ? S - k0 : ACC_CHUNK;
                const uint32_t run = c * SLOT_GROUP;  // a group's slots for samples k0 .. k0 + c - 1
#pragma unroll
                for (uint32_t target = 0; target < 4; ++target) {
                    if (target > 0 && !aov) break;
                    auto load = [&](uint32_t slot) {
                        if (target == 0) return tg.radiance[slot];
                        if (target == 1) return tg.albedo[slot];
                        if (target == 2) return tg.normal[slot];
                        return make_float4(tg.depth[slot], 0.0f, 0.0f, 0.0f);
                    };
                    // Element e of the chunk: group e / run, offset e % run in the group's run.
                    auto place = [&](uint32_t e, uint32_t q, uint32_t off) -> float4& {
                        return stage[(q * SLOT_GROUP + off % SLOT_GROUP) * (ACC_CHUNK + 1) + off / SLOT_GROUP];
                    };
                    if (c == ACC_CHUNK) {
                        // Every load of the pass in flight before the first LDS store.
                        constexpr uint32_t RUN = ACC_CHUNK * SLOT_GROUP;
                        float4 v[ACC_CHUNK];
#pragma unroll
                        for (uint32_t j = 0; j < ACC_CHUNK; ++j) {
                            const uint32_t e = threadIdx.x + j * BLOCK, q = e / RUN, off = e % RUN;
                            v[j] = load(tile_slot + q * SLOT_GROUP * S + k0 * SLOT_GROUP + off);
                        }
#pragma unroll
                        for (uint32_t j = 0; j < ACC_CHUNK; ++j) {
                            const uint32_t e = threadIdx.x + j * BLOCK, q = e / RUN, off = e % RUN;
                            place(e, q, off) = v[j];
                        }
                    } else {
                        for (uint32_t e = threadIdx.x; e < c * BLOCK; e += BLOCK) {
                            const uint32_t q = e / run, off = e - q * run;
                            place(e, q, off) = load(tile_slot + q * SLOT_GROUP * S + k0 * SLOT_GROUP + off);
                        }
                    }
                    __syncthreads();
                    const float4* mine = stage + threadIdx.x * (ACC_CHUNK + 1);
                    for (uint32_t kk = 0; kk < c; ++kk) {
                        if (target == 0) sums.radiance(k0 + kk, mine[kk]);
                        else if (target == 1) sums.albedo(k0 + kk, mine[kk]);
                        else if (target == 2) sums.normal(k0 + kk, mine[kk]);
                        else sums.depth(k0 + kk, mine[kk].x);
                    }
                    __syncthreads();
                }
            }
            store_sums(job, px, sums);
            continue;
        }
        const uint32_t i = tile + threadIdx.x;
        if (i >= total_pixels) continue;
        const JobRecord& job = jobs[job_of(jobs, job_count, i)];
        if (!job.beauty_sum) continue;
        // px is the film pixel that band-order index holds; the pixel's
        // slots are job_slot(index, k) = slot0 + k * stride.
        const uint32_t index = i - job.pixel_base;
        const uint32_t px = band_pixel(index, job.width, job.height);
        const bool aov = tg.albedo != nullptr && job.albedo_sum != nullptr;
        PixelSums sums = load_sums(job, px, aov);
        const uint32_t slot0 = job.slot_base + job_slot(index, 0, job.pixels, job.samples);
        const uint32_t stride = job_slot(index, 1, job.pixels, job.samples) - job_slot(index, 0, job.pixels, job.samples);
        for (uint32_t k = 0; k < job.samples; ++k) {
            const uint32_t slot = slot0 + k * stride;
            sums.radiance(k, tg.radiance[slot]);
            if (aov) {
                sums.albedo(k, tg.albedo[slot]);
                sums.normal(k, tg.normal[slot]);
                sums.depth(k, tg.depth[slot]);
            }
        }
        store_sums(job, px, sums);
    }
}

__global__ __launch_bounds__(BLOCK) void k_resolve(FilmView f, uint32_t pixels, float samples) {
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < pixels; i += gridDim.x * BLOCK) {
        const float4 b = f.beauty_sum[i], a = f.albedo_sum[i], nn = f.normal_sum[i];
        f.beauty[i] = make_float4(b.x / samples, b.y / samples, b.z / samples, b.w / samples);
        f.albedo[i] = make_float4(a.x / samples, a.y / samples, a.z / samples, a.w / samples);
        f.normal[i] = make_float4(nn.x / samples, nn.y / samples, nn.z / samples, nn.w / samples);
        f.depth[i] = f.depth_sum[i] / samples;
    }
}

// The estimated relative MSE of a film's running mean, for noise-target
// stopping: per pixel, the variance of the mean luminance over its squared
// mean plus 1e-2 (the relMSE denominator), from the luminance sum and the
// squared-luminance sum; one f64 partial sum per block, summed on the host.
__global__ __launch_bounds__(BLOCK) void k_noise(const float4* __restrict__ sums, const float* __restrict__ moments,
                                                 uint32_t pixels, float samples, double* partials) {
    __shared__ double lane_sums[BLOCK];
    double acc = 0.0;
    const double n = samples;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < pixels; i += gridDim.x * BLOCK) {
        const float4 b = sums[i];
        const double mean = ((double(b.x) * LUMA_R + double(b.y) * LUMA_G) + double(b.z) * LUMA_B) / n;
        const double second = double(moments[i]) / n;
        const double var_mean = fmax(0.0, second - mean * mean) / (n - 1.0);
        acc += var_mean / (mean * mean + 1e-2);
    }
    lane_sums[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t w = BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) lane_sums[threadIdx.x] += lane_sums[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = lane_sums[0];
}

// --- diagnostics ----------------------------------------------------------------------

// rng_test.slang:21-29: the first `count` points of one (pixel, dimension) key.
__global__ void k_sample_sequence(uint32_t pixel, uint32_t dimension, uint32_t count, float* points, float* values) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const f2 pt = sample_2d(pixel, i, dimension);
    points[2 * i] = pt.x;
    points[2 * i + 1] = pt.y;
    values[i] = sample_1d(pixel, i, dimension);
}

// The bounce-0 closest hit per pixel, for the first-hit identity check.
template <bool W8>
__global__ __launch_bounds__(TRAVERSAL_BLOCK) void k_first_hits(SceneView s, CameraView cam, PathPool p, uint32_t width,
                                                      uint32_t height, uint32_t sample, uint32_t* inst_out,
                                                      uint32_t* prim_out, float* bary_out) {
    __shared__ uint32_t lds_stack[STACK_LDS * TRAVERSAL_BLOCK];
    const uint32_t n = width * height;
    const uint32_t lanes = gridDim.x * TRAVERSAL_BLOCK;
    Stack st = make_stack(lds_stack, p.stack_overflow, lanes, blockIdx.x * TRAVERSAL_BLOCK + threadIdx.x);
    uint32_t nodes = 0, tris = 0;
    for (uint32_t first = blockIdx.x * TRAVERSAL_BLOCK; first < n; first += lanes) {
        const uint32_t pixel = first + threadIdx.x;
        if (pixel >= n) continue;
        v3 o, d;
        float cf;
        camera_ray(cam, width, height, pixel, sample, &o, &d, &cf);
        const TraceHit h = trace_closest<W8>(s, o, d, 1e30f, MASK_CAMERA, st, pixel, sample, 0, nodes, tris);
        inst_out[pixel] = h.hit ? h.inst : 0xffffffffu;
        prim_out[pixel] = h.hit ? h.prim : 0xffffffffu;
        bary_out[2 * pixel] = h.hit ? h.b1 : 0.0f;
        bary_out[2 * pixel + 1] = h.hit ? h.b2 : 0.0f;
    }
}

// Environment importance sampling for given uniforms (the CDF inversion
// check): per sample, the direction, pdf and radiance of env_sample and the
// (v, u) position and (row, column) the two inversions chose.
__global__ __launch_bounds__(BLOCK) void k_env_sample(SceneView s, const float2* __restrict__ u, uint32_t n,
                                                      float* __restrict__ out) {
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        const float2 ui = u[i];
        uint32_t row, col;
        float v, uu;
        sample_cdf(s.env_marginal, s.env_marginal_guide, s.env_marginal_buckets, s.env_height, ui.x, &row, &v);
        sample_cdf(s.env_conditional + row * (s.env_width + 1),
                   s.env_conditional_guide + row * (s.env_conditional_buckets + 1), s.env_conditional_buckets,
                   s.env_width, ui.y, &col, &uu);
        const EnvSample e = env_sample(s, {ui.x, ui.y});
        float* o = out + 12 * size_t(i);
        o[0] = e.direction.x;
        o[1] = e.direction.y;
        o[2] = e.direction.z;
        o[3] = e.pdf;
        o[4] = e.radiance.x;
        o[5] = e.radiance.y;
        o[6] = e.radiance.z;
        o[7] = v;
        o[8] = uu;
        o[9] = __uint_as_float(row);
        o[10] = __uint_as_float(col);
        o[11] = 0.0f;
    }
}

// --- launchers --------------------------------------------------------------------------

static uint32_t grid_for(uint32_t items, uint32_t cap) {
    uint32_t blocks = (items + BLOCK - 1) / BLOCK;
    if (blocks < 1) blocks = 1;
    return blocks < cap ? blocks : cap;
}

hipError_t launch_raygen(hipStream_t stream, uint32_t* counters, uint32_t counter_lines, Profiler* prof) {
    if (prof) prof->begin(K_RAYGEN);
    hipLaunchKernelGGL(k_raygen, dim3((counter_lines + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream, counters,
                       counter_lines);
    if (prof) prof->end(K_RAYGEN);
    return hipGetLastError();
}

hipError_t launch_bounce(hipStream_t stream, const LaunchShape& shape, const SceneView& s, const PathPool& p,
                         const WaveTargets& tg, uint32_t* counters, const JobRecord* jobs, uint32_t job_count,
                         uint32_t base, uint32_t range_count, uint32_t bounce, uint32_t light_sampling, int counting,
                         bool tail, uint32_t* tail_claims, uint32_t compute_units, Profiler* prof) {
    const uint32_t trace_grid = shape.trace_blocks;
    const uint32_t light_grid = grid_for(range_count, shape.light_blocks);
    // Bounce b's rays and bounce b-1's shadow rays share one launch; the
    // last bounce shades terminally and leaves no shadow rays behind.
    if (prof) prof->begin(K_TRACE);
    const bool wide = shape.bvh_width == 8 && s.nodes8 != nullptr;
#define CN_TRACE(COUNT, W8, CAMERA)                                                                                 \
    hipLaunchKernelGGL((k_trace<COUNT, W8, CAMERA>), dim3(trace_grid), dim3(TRAVERSAL_BLOCK), 0, stream, s, p, tg,   \
                       counters, jobs, job_count, base, range_count, bounce, light_sampling)
#define CN_TRACE_CAMERA(COUNT, W8)   \
    if (bounce == 0) CN_TRACE(COUNT, W8, true); \
    else CN_TRACE(COUNT, W8, false)
    if (counting) {
        if (wide) CN_TRACE_CAMERA(true, true);
        else CN_TRACE_CAMERA(true, false);
    } else {
        if (wide) CN_TRACE_CAMERA(false, true);
        else CN_TRACE_CAMERA(false, false);
    }
#undef CN_TRACE_CAMERA
#undef CN_TRACE
    if (prof) prof->end(K_TRACE);
    if (tail) {
        // The waves the chip holds at once (TAIL_WAVES per SIMD: one
        // 256-lane block per CU for each), at most the trace grid (whose
        // lanes the traversal stack's spill area is sized for): the pairs
        // claim items until none are left.
        const uint32_t resident = (compute_units ? compute_units : 1u) * TAIL_WAVES;
        const uint32_t tail_grid = resident < trace_grid ? resident : trace_grid;
        if (prof) prof->begin(K_TAIL);
        if (wide)
            hipLaunchKernelGGL(k_tail<true>, dim3(tail_grid), dim3(TRAVERSAL_BLOCK), 0, stream, s, p, tg, counters,
                               tail_claims, base, bounce, light_sampling);
        else
            hipLaunchKernelGGL(k_tail<false>, dim3(tail_grid), dim3(TRAVERSAL_BLOCK), 0, stream, s, p, tg, counters,
                               tail_claims, base, bounce, light_sampling);
        if (prof) prof->end(K_TAIL);
        return hipGetLastError();
    }
    if (bounce > 0) {  // bounce 0's misses were shaded by the trace launch
        if (prof) prof->begin(K_SHADE_MISS);
        hipLaunchKernelGGL(k_shade_miss, dim3(light_grid), dim3(BLOCK), 0, stream, s, p, tg, counters, base, bounce,
                           light_sampling);
        if (prof) prof->end(K_SHADE_MISS);
    }
    if (prof) prof->begin(K_SHADE_SURFACE);
    const uint32_t sb = shade_block(shape.shade_variant);
    uint32_t shade_grid = (range_count + sb - 1) / sb;
    if (shade_grid > shape.shade_blocks) shade_grid = shape.shade_blocks;
    if (shade_grid < 1) shade_grid = 1;
#define CN_SHADE_LAUNCH(SB, LDS, SORT)                                                                      \
    hipLaunchKernelGGL((k_shade_surface<SB, LDS, SORT>), dim3(shade_grid), dim3(SB), 0, stream, s, p, tg, counters, \
                       base, bounce, light_sampling)
    switch (shape.shade_variant) {
        case SHADE_1024_LDS: CN_SHADE_LAUNCH(1024, true, false); break;
        case SHADE_512_LDS: CN_SHADE_LAUNCH(512, true, false); break;
        case SHADE_256_LDS: CN_SHADE_LAUNCH(256, true, false); break;
        case SHADE_256_GLOBAL: CN_SHADE_LAUNCH(256, false, false); break;
        case SHADE_512_LDS_SORTED: CN_SHADE_LAUNCH(512, true, true); break;
        case SHADE_256_LDS_SORTED: CN_SHADE_LAUNCH(256, true, true); break;
        default: CN_SHADE_LAUNCH(256, false, true); break;
    }
#undef CN_SHADE_LAUNCH
    if (prof) prof->end(K_SHADE_SURFACE);
    return hipGetLastError();
}

hipError_t launch_accumulate(hipStream_t stream, const JobRecord* jobs, uint32_t job_count, const WaveTargets& tg,
                             uint32_t total_pixels, Profiler* prof) {
    if (prof) prof->begin(K_ACCUMULATE);
    hipLaunchKernelGGL(k_accumulate, dim3(grid_for(total_pixels, 4096)), dim3(BLOCK), 0, stream, jobs, job_count, tg,
                       total_pixels);
    if (prof) prof->end(K_ACCUMULATE);
    return hipGetLastError();
}

hipError_t launch_resolve(hipStream_t stream, const FilmView& f, uint32_t pixels, float samples) {
    hipLaunchKernelGGL(k_resolve, dim3(grid_for(pixels, 4096)), dim3(BLOCK), 0, stream, f, pixels, samples);
    return hipGetLastError();
}

hipError_t launch_noise(hipStream_t stream, const float4* sums, const float* moments, uint32_t pixels, float samples,
                        double* partials, uint32_t blocks) {
    hipLaunchKernelGGL(k_noise, dim3(blocks), dim3(BLOCK), 0, stream, sums, moments, pixels, samples, partials);
    return hipGetLastError();
}

hipError_t launch_sample_sequence(hipStream_t stream, uint32_t pixel, uint32_t dimension, uint32_t count,
                                  float* points, float* values) {
    hipLaunchKernelGGL(k_sample_sequence, dim3((count + 255) / 256), dim3(256), 0, stream, pixel, dimension, count,
                       points, values);
    return hipGetLastError();
}

hipError_t launch_first_hits(hipStream_t stream, const LaunchShape& shape, const SceneView& s, const CameraView& cam,
                             const PathPool& p, uint32_t width, uint32_t height, uint32_t sample, uint32_t* inst,
                             uint32_t* prim, float* bary) {
    if (shape.bvh_width == 8 && s.nodes8 != nullptr)
        hipLaunchKernelGGL(k_first_hits<true>, dim3(shape.trace_blocks), dim3(TRAVERSAL_BLOCK), 0, stream, s, cam, p,
                           width, height, sample, inst, prim, bary);
    else
        hipLaunchKernelGGL(k_first_hits<false>, dim3(shape.trace_blocks), dim3(TRAVERSAL_BLOCK), 0, stream, s, cam, p,
                           width, height, sample, inst, prim, bary);
    return hipGetLastError();
}

hipError_t launch_env_sample(hipStream_t stream, const SceneView& s, const float2* u, uint32_t n, float* out) {
    hipLaunchKernelGGL(k_env_sample, dim3(grid_for(n, 1024)), dim3(BLOCK), 0, stream, s, u, n, out);
    return hipGetLastError();
}

}  // namespace cn
