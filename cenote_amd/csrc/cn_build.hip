// cn_build.hip — the BVH built on the device (the reference builds its
// acceleration structures with the driver, gpu/accel.rs:167-269).
//
// The binary tree is the host builder's (cn_bvh.cpp) bit for bit: every
// split is made by bvh_decide's rule from bins that are exact in any order
// (min/max folds, integer counts), and ranges are split by a stable
// partition. Everything runs on the device; the host only launches and, once
// every eight levels, reads how many large ranges are left:
//
//   * top levels, ranges of more than SMALL_TASK triangles, one level at a
//     time: device-wide passes over 8192-triangle chunks (bounds, then
//     32-bin histograms on three axes), each range's split decided by one
//     workgroup from those device-resident bins, then a stable scatter
//     (per-chunk counts, per-range running sums, in-chunk scans);
//   * every remaining range: one 256-lane workgroup builds its whole subtree
//     in LDS, depth first, smaller child first;
//   * the collapse to 4-wide and 8-wide nodes, in the host collapse's
//     preorder numbering, and their quantization.
//
// The result equals the host builder's: the same binary tree, the same
// full-precision 4-wide nodes, the same quantized 4- and 8-wide nodes and
// the same leaf-order triangles.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "cn_bvh.h"

namespace cn {

namespace {

constexpr uint32_t BUILD_BLOCK = 256;
constexpr uint32_t CHUNK = 8192;       // triangles per workgroup in the device-wide passes
constexpr uint32_t SMALL_TASK = 2048;  // ranges at most this large are built by one workgroup
constexpr uint32_t SMALL_STACK = 32;   // ample: smaller-first keeps it under log2(SMALL_TASK) + 1

// Floats as order-preserving unsigned integers, so integer atomic min/max
// fold float bounds exactly.
__device__ __forceinline__ uint32_t f2o(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float o2f(uint32_t o) {
    const uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
constexpr uint32_t O_POS_INF = 0xff800000u;  // f2o(+inf): the identity of min
constexpr uint32_t O_NEG_INF = 0x007fffffu;  // f2o(-inf): the identity of max

// Per-triangle bounds and centroids as structure-of-arrays.
struct PrimArrays {
    float* lo[3];
    float* hi[3];
    float* c[3];
};

struct Chunk {
    uint32_t task, begin, end, pad;
};

// A top-level range's split, as the chunk passes read it.
struct TaskSplit {
    uint32_t begin, count;  // the range
    uint32_t mode;          // 0: by bin (axis, split); 1: by index (first `left` go left)
    uint32_t axis;
    int32_t split;
    uint32_t left;
    float lo, scale;        // the axis's centroid low bound and bin scale
};

// Bounds of a task: box lo[3], box hi[3], centroid lo[3], centroid hi[3].
constexpr uint32_t TB = 12;
// Bins of a task: 3 axes x 32 bins x (count + box lo[3] + box hi[3]).
constexpr uint32_t TBIN = 3 * BVH_BINS * 7;

__global__ __launch_bounds__(BUILD_BLOCK) void k_prims(const WorldTriangle* __restrict__ in, uint32_t n,
                                                       PrimArrays p, uint32_t* order, uint32_t* max_abs_bits,
                                                       uint32_t* bad) {
    const uint32_t i = blockIdx.x * BUILD_BLOCK + threadIdx.x;
    if (i >= n) return;
    const WorldTriangle t = in[i];
    const float v[3][3] = {{t.v0.x, t.v0.y, t.v0.z}, {t.v1.x, t.v1.y, t.v1.z}, {t.v2.x, t.v2.y, t.v2.z}};
    Box3 box;
    float c[3];
    bvh_prim_bounds(v, &box, c);
    float m = 0.0f;
    bool finite = true;
    for (int q = 0; q < 3; ++q)
        for (int k = 0; k < 3; ++k) {
            m = fmaxf(m, fabsf(v[q][k]));
            finite = finite && isfinite(v[q][k]);
        }
    for (int k = 0; k < 3; ++k) {
        p.lo[k][i] = box.lo[k];
        p.hi[k][i] = box.hi[k];
        p.c[k][i] = c[k];
    }
    order[i] = i;
    atomicMax(max_abs_bits, __float_as_uint(m));  // non-negative floats order as their bits
    if (!finite) atomicOr(bad, 1u);
}

// 32-bin histograms on the three axes of one chunk, folded into its task's.
__device__ void k_chunk_bins_body(const Chunk& ch, const PrimArrays& p, const uint32_t* __restrict__ order,
                                  const uint32_t* __restrict__ tb, uint32_t* bins) {
    __shared__ uint32_t lb[TBIN];
    for (uint32_t k = threadIdx.x; k < TBIN; k += BUILD_BLOCK) {
        const uint32_t f = k % 7;
        lb[k] = f == 0 ? 0u : (f <= 3 ? O_POS_INF : O_NEG_INF);
    }
    __syncthreads();
    const uint32_t* b = tb + size_t(ch.task) * TB;
    float lo[3], scale[3];
    bool live[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = o2f(b[6 + k]);
        const float hi = o2f(b[9 + k]);
        live[k] = hi - lo[k] > 0.0f;
        scale[k] = live[k] ? bvh_scale(lo[k], hi) : 0.0f;
    }
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += BUILD_BLOCK) {
        const uint32_t id = order[i];
        uint32_t box[6];
        for (int k = 0; k < 3; ++k) {
            box[k] = f2o(p.lo[k][id]);
            box[3 + k] = f2o(p.hi[k][id]);
        }
        for (int axis = 0; axis < 3; ++axis) {
            if (!live[axis]) continue;
            const int bin = bvh_bin(p.c[axis][id], lo[axis], scale[axis]);
            uint32_t* e = &lb[(axis * BVH_BINS + bin) * 7];
            atomicAdd(&e[0], 1u);
            for (int k = 0; k < 3; ++k) {
                atomicMin(&e[1 + k], box[k]);
                atomicMax(&e[4 + k], box[3 + k]);
            }
        }
    }
    __syncthreads();
    uint32_t* g = bins + size_t(ch.task) * TBIN;
    for (uint32_t k = threadIdx.x; k < TBIN; k += BUILD_BLOCK) {
        const uint32_t f = k % 7;
        if (f == 0) {
            if (lb[k]) atomicAdd(&g[k], lb[k]);
        } else if (f <= 3) {
            atomicMin(&g[k], lb[k]);
        } else {
            atomicMax(&g[k], lb[k]);
        }
    }
}

__device__ __forceinline__ bool goes_left(const TaskSplit& s, uint32_t i, uint32_t id, const PrimArrays& p) {
    if (s.mode == 1) return i - s.begin < s.left;
    return bvh_bin(p.c[s.axis][id], s.lo, s.scale) <= s.split;
}

// Stable scatter of one chunk: lanes take contiguous runs, an exclusive
// scan of the runs' left counts orders them.
__device__ void k_chunk_scatter_body(const Chunk& ch, const TaskSplit* __restrict__ splits, const PrimArrays& p,
                                     const uint32_t* __restrict__ order, uint32_t left_base, uint32_t right_base,
                                     uint32_t* out) {
    __shared__ uint32_t scan[BUILD_BLOCK];
    const TaskSplit s = splits[ch.task];
    const uint32_t len = ch.end - ch.begin;
    const uint32_t per = (len + BUILD_BLOCK - 1) / BUILD_BLOCK;
    const uint32_t r0 = min(ch.begin + threadIdx.x * per, ch.end);
    const uint32_t r1 = min(r0 + per, ch.end);
    uint32_t mine = 0;
    for (uint32_t i = r0; i < r1; ++i) mine += goes_left(s, i, order[i], p);
    scan[threadIdx.x] = mine;
    __syncthreads();
    for (uint32_t off = 1; off < BUILD_BLOCK; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? scan[threadIdx.x - off] : 0u;
        __syncthreads();
        scan[threadIdx.x] += add;
        __syncthreads();
    }
    const uint32_t lefts_before = scan[threadIdx.x] - mine;
    const uint32_t rights_before = (r0 - ch.begin) - lefts_before;
    uint32_t l = left_base + lefts_before;
    uint32_t r = right_base + rights_before;
    for (uint32_t i = r0; i < r1; ++i) {
        const uint32_t id = order[i];
        if (goes_left(s, i, id, p)) out[l++] = id;
        else out[r++] = id;
    }
}


// --- the split decision, taken by a whole workgroup ----------------------------

// bvh_decide (cn_bvh_split.h) with its SAH sweep spread over lanes: lane
// axis * 32 + b (b < 31) prices the split after bin b of `axis` from the
// bins' min/max folds and integer counts — the very boxes, areas and costs
// the sequential sweep computes, since a fold of boxes is the same in any
// order — and a (cost, candidate index) minimum picks the candidate the
// sequential sweep keeps (the first of the lowest cost). Thread 0 returns
// the decision (with the left count) in `*out`. Every thread of the block
// must call it.
struct Decision {
    SplitDecision d;
    uint32_t left;  // primitives in bins 0..split of the chosen axis
};

__device__ void decide_block(const Box3 (*fbins)[BVH_BINS], const uint32_t (*fcounts)[BVH_BINS], const Box3& box,
                             const Box3& cbox, uint32_t n, float* red_cost, uint32_t* red_idx, uint32_t* red_left,
                             Decision* out) {
    const uint32_t tid = threadIdx.x;
    float cost = INFINITY;
    uint32_t idx = 0xffffffffu, lcnt = 0;
    if (tid < 3 * BVH_BINS && n > 1) {
        const int axis = int(tid / BVH_BINS), b = int(tid % BVH_BINS);
        if (b < BVH_BINS - 1 && cbox.hi[axis] - cbox.lo[axis] > 0.0f) {
            Box3 lacc = Box3::empty(), racc = Box3::empty();
            uint32_t rcnt = 0;
            for (int k = 0; k < BVH_BINS; ++k) {
                if (k <= b) {
                    lacc.grow(fbins[axis][k]);
                    lcnt += fcounts[axis][k];
                } else {
                    racc.grow(fbins[axis][k]);
                    rcnt += fcounts[axis][k];
                }
            }
            if (lcnt != 0 && rcnt != 0) {
                cost = lacc.area() * float(lcnt) + racc.area() * float(rcnt);
                idx = tid;
            }
        }
    }
    // Minimum of (cost, idx) over each wave, then over the two waves that hold candidates.
    for (int offset = 32; offset > 0; offset >>= 1) {
        const float oc = __shfl_xor(cost, offset);
        const uint32_t oi = __shfl_xor(idx, offset);
        const uint32_t ol = __shfl_xor(lcnt, offset);
        if (oc < cost || (oc == cost && oi < idx)) {
            cost = oc;
            idx = oi;
            lcnt = ol;
        }
    }
    const uint32_t wave = tid / 64;
    if ((tid & 63) == 0 && wave < 2) {
        red_cost[wave] = cost;
        red_idx[wave] = idx;
        red_left[wave] = lcnt;
    }
    __syncthreads();
    if (tid == 0) {
        float best = red_cost[0];
        uint32_t bi = red_idx[0], bl = red_left[0];
        if (red_cost[1] < best || (red_cost[1] == best && red_idx[1] < bi)) {
            best = red_cost[1];
            bi = red_idx[1];
            bl = red_left[1];
        }
        Decision r;
        r.d = SplitDecision{0, -1, 0};
        r.left = 0;
        if (n <= 1) {
            r.d.leaf = 1;
        } else {
            if (bi != 0xffffffffu) {
                r.d.axis = int(bi / BVH_BINS);
                r.d.split = int(bi % BVH_BINS);
                r.left = bl;
            }
            const float parent_area = box.area();
            const float leaf_cost = float(n);
            const float split_cost = BVH_NODE_COST + (parent_area > 0.0f ? best / parent_area : 0.0f);
            if (n <= BVH_LEAF_SIZE && (r.d.axis < 0 || leaf_cost <= split_cost)) r.d.leaf = 1;
        }
        *out = r;
    }
    __syncthreads();
}

// Block-wide min/max fold of 12 ordered-integer bounds (box lo, box hi,
// centroid lo, centroid hi) into red[0..11]: a wave reduction by shuffles,
// then one word per wave through LDS.
__device__ void fold_bounds(uint32_t (&v)[TB], uint32_t* red, uint32_t* wave_red) {
#pragma unroll
    for (int k = 0; k < int(TB); ++k) {
        uint32_t x = v[k];
        for (int offset = 32; offset > 0; offset >>= 1) {
            const uint32_t o = __shfl_xor(x, offset);
            x = (k % 6) < 3 ? min(x, o) : max(x, o);
        }
        v[k] = x;
    }
    const uint32_t wave = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < int(TB); ++k) wave_red[wave * TB + k] = v[k];
    __syncthreads();
    if (threadIdx.x < TB) {
        const uint32_t k = threadIdx.x;
        uint32_t x = wave_red[k];
        for (uint32_t w = 1; w < BUILD_BLOCK / 64; ++w)
            x = (k % 6) < 3 ? min(x, wave_red[w * TB + k]) : max(x, wave_red[w * TB + k]);
        red[k] = x;
    }
    __syncthreads();
}

// --- top levels: ranges of more than SMALL_TASK triangles ----------------------

struct Task {
    uint32_t begin, end, node, pad;
};

// Build-state words on the device.
enum : uint32_t {
    S_MAX_ABS = 0,   // the largest |coordinate| as float bits
    S_BAD = 1,       // a non-finite vertex
    S_NODES = 2,     // binary nodes allocated
    S_OVERFLOW = 3,  // a subtree stack or a task list overflowed
    S_BIG = 4,       // big tasks in the current level's list
    S_NEXT_BIG = 5,  // big tasks appended for the next level
    S_SMALL = 6,     // small tasks appended
    S_CHUNKS = 7,    // chunks of the current level
    S_WIDE = 8,      // wide nodes allocated (collapse)
    S_LIST = 9,      // collapse list entries appended
    S_WORDS = 10
};

// One level's chunk list (per big task, CHUNK-triangle chunks in range
// order), each task's first chunk, and the identities its bounds and bins
// fold into. One workgroup: the tasks' chunk counts are prefix-summed by
// the workgroup's lanes, each taking a contiguous run of tasks.
__global__ __launch_bounds__(BUILD_BLOCK) void k_level_chunks(const Task* __restrict__ tasks, uint32_t* state,
                                                              Chunk* chunks, uint32_t* chunk_first,
                                                              uint32_t chunk_capacity) {
    __shared__ uint32_t scan[BUILD_BLOCK];
    const uint32_t nt = state[S_BIG];
    const uint32_t per = (nt + BUILD_BLOCK - 1) / BUILD_BLOCK;
    const uint32_t t0 = min(threadIdx.x * per, nt), t1 = min(t0 + per, nt);
    uint32_t mine = 0;
    for (uint32_t t = t0; t < t1; ++t) mine += (tasks[t].end - tasks[t].begin + CHUNK - 1) / CHUNK;
    scan[threadIdx.x] = mine;
    __syncthreads();
    for (uint32_t off = 1; off < BUILD_BLOCK; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? scan[threadIdx.x - off] : 0u;
        __syncthreads();
        scan[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t at = scan[threadIdx.x] - mine;
    for (uint32_t t = t0; t < t1; ++t) {
        chunk_first[t] = at;
        for (uint32_t b = tasks[t].begin; b < tasks[t].end; b += CHUNK) {
            if (at < chunk_capacity) chunks[at] = Chunk{t, b, min(b + CHUNK, tasks[t].end), 0};
            ++at;
        }
    }
    if (threadIdx.x == BUILD_BLOCK - 1) {
        chunk_first[nt] = scan[BUILD_BLOCK - 1];
        state[S_CHUNKS] = min(scan[BUILD_BLOCK - 1], chunk_capacity);
        if (scan[BUILD_BLOCK - 1] > chunk_capacity) atomicOr(&state[S_OVERFLOW], 1u);
        state[S_NEXT_BIG] = 0;
    }
}

// Identities for every big task's bounds (min for lo words, max for hi) and
// bins (count 0, box lo min identity, box hi max identity).
__global__ __launch_bounds__(BUILD_BLOCK) void k_level_init(const uint32_t* __restrict__ state, uint32_t* tb,
                                                            uint32_t* bins) {
    const uint32_t nt = state[S_BIG];
    for (uint32_t i = blockIdx.x * BUILD_BLOCK + threadIdx.x; i < nt * TBIN; i += gridDim.x * BUILD_BLOCK) {
        const uint32_t f = i % 7;
        bins[i] = f == 0 ? 0u : (f <= 3 ? O_POS_INF : O_NEG_INF);
        if (i < nt * TB) tb[i] = (i % 6) < 3 ? O_POS_INF : O_NEG_INF;
    }
}

// The chunk passes read their chunk from the level's list; blocks past its
// end have nothing to do.
__global__ __launch_bounds__(BUILD_BLOCK) void k_level_bounds(const Chunk* __restrict__ chunks,
                                                              const uint32_t* __restrict__ state, PrimArrays p,
                                                              const uint32_t* __restrict__ order, uint32_t* tb) {
    if (blockIdx.x >= state[S_CHUNKS]) return;
    __shared__ uint32_t red[TB];
    const Chunk ch = chunks[blockIdx.x];
    if (threadIdx.x < TB) red[threadIdx.x] = (threadIdx.x % 6) < 3 ? O_POS_INF : O_NEG_INF;
    __syncthreads();
    uint32_t v[TB];
    for (int k = 0; k < int(TB); ++k) v[k] = (k % 6) < 3 ? O_POS_INF : O_NEG_INF;
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += BUILD_BLOCK) {
        const uint32_t id = order[i];
        for (int k = 0; k < 3; ++k) {
            v[k] = min(v[k], f2o(p.lo[k][id]));
            v[3 + k] = max(v[3 + k], f2o(p.hi[k][id]));
            const uint32_t c = f2o(p.c[k][id]);
            v[6 + k] = min(v[6 + k], c);
            v[9 + k] = max(v[9 + k], c);
        }
    }
    for (int k = 0; k < int(TB); ++k) {
        uint32_t x = v[k];
        for (int offset = 32; offset > 0; offset >>= 1) {
            const uint32_t o = __shfl_xor(x, offset);
            x = (k % 6) < 3 ? min(x, o) : max(x, o);
        }
        if ((threadIdx.x & 63) == 0) {
            if ((k % 6) < 3) atomicMin(&red[k], x);
            else atomicMax(&red[k], x);
        }
    }
    __syncthreads();
    if (threadIdx.x < TB) {
        uint32_t* dst = tb + size_t(ch.task) * TB + threadIdx.x;
        if ((threadIdx.x % 6) < 3) atomicMin(dst, red[threadIdx.x]);
        else atomicMax(dst, red[threadIdx.x]);
    }
}

__global__ __launch_bounds__(BUILD_BLOCK) void k_level_bins(const Chunk* __restrict__ chunks,
                                                            const uint32_t* __restrict__ state, PrimArrays p,
                                                            const uint32_t* __restrict__ order,
                                                            const uint32_t* __restrict__ tb, uint32_t* bins) {
    if (blockIdx.x >= state[S_CHUNKS]) return;
    k_chunk_bins_body(chunks[blockIdx.x], p, order, tb, bins);
}

// Each big task's split, decided by one workgroup from its device-resident
// bounds and bins (the host builder's bvh_decide, in parallel): the task's
// binary node gets its box and two fresh children, and each child range
// goes to the next level's big list or to the small list.
__global__ __launch_bounds__(BUILD_BLOCK) void k_level_decide(const Task* __restrict__ tasks, uint32_t* state,
                                                              const uint32_t* __restrict__ tb,
                                                              const uint32_t* __restrict__ bins, TaskSplit* splits,
                                                              BinaryNode* nodes, Task* next_big, Task* small,
                                                              uint32_t big_capacity, uint32_t small_capacity) {
    __shared__ Box3 fbins[3][BVH_BINS];
    __shared__ uint32_t fcounts[3][BVH_BINS];
    __shared__ float red_cost[2];
    __shared__ uint32_t red_idx[2], red_left[2];
    __shared__ Decision dec;
    if (blockIdx.x >= state[S_BIG]) return;
    const uint32_t k = blockIdx.x;
    const Task task = tasks[k];
    const uint32_t m = task.end - task.begin;
    const uint32_t* t = tb + size_t(k) * TB;
    Box3 box, cbox;
    for (int a = 0; a < 3; ++a) {
        box.lo[a] = o2f(t[a]);
        box.hi[a] = o2f(t[3 + a]);
        cbox.lo[a] = o2f(t[6 + a]);
        cbox.hi[a] = o2f(t[9 + a]);
    }
    if (threadIdx.x < 3 * BVH_BINS) {
        const uint32_t a = threadIdx.x / BVH_BINS, b = threadIdx.x % BVH_BINS;
        const bool live = cbox.hi[a] - cbox.lo[a] > 0.0f;
        const uint32_t* e = bins + size_t(k) * TBIN + threadIdx.x * 7;
        fcounts[a][b] = live ? e[0] : 0u;
        for (int q = 0; q < 3; ++q) {
            fbins[a][b].lo[q] = live ? o2f(e[1 + q]) : INFINITY;
            fbins[a][b].hi[q] = live ? o2f(e[4 + q]) : -INFINITY;
        }
    }
    __syncthreads();
    decide_block(fbins, fcounts, box, cbox, m, red_cost, red_idx, red_left, &dec);
    if (threadIdx.x != 0) return;
    TaskSplit sp{task.begin, m, 1u, 0u, 0, m / 2, 0.0f, 0.0f};
    if (dec.d.axis >= 0 && dec.left > 0 && dec.left < m) {
        sp.mode = 0;
        sp.axis = uint32_t(dec.d.axis);
        sp.split = dec.d.split;
        sp.left = dec.left;
        sp.lo = cbox.lo[dec.d.axis];
        sp.scale = bvh_scale(cbox.lo[dec.d.axis], cbox.hi[dec.d.axis]);
    }
    splits[k] = sp;
    // Ranges this large always split (a leaf holds at most BVH_LEAF_SIZE triangles).
    const uint32_t ids = atomicAdd(&state[S_NODES], 2u);
    BinaryNode node;
    node.box = box;
    node.left = ids;
    node.right = ids + 1;
    node.first = node.count = 0;
    nodes[task.node] = node;
    const Task kids[2] = {{task.begin, task.begin + sp.left, ids, 0}, {task.begin + sp.left, task.end, ids + 1, 0}};
    for (const Task& kid : kids) {
        if (kid.end - kid.begin > SMALL_TASK) {
            const uint32_t at = atomicAdd(&state[S_NEXT_BIG], 1u);
            if (at < big_capacity) next_big[at] = kid;
            else atomicOr(&state[S_OVERFLOW], 2u);
        } else {
            const uint32_t at = atomicAdd(&state[S_SMALL], 1u);
            if (at < small_capacity) small[at] = kid;
            else atomicOr(&state[S_OVERFLOW], 4u);
        }
    }
}

__global__ __launch_bounds__(BUILD_BLOCK) void k_level_count(const Chunk* __restrict__ chunks,
                                                             const uint32_t* __restrict__ state,
                                                             const TaskSplit* __restrict__ splits, PrimArrays p,
                                                             const uint32_t* __restrict__ order, uint32_t* lefts) {
    if (blockIdx.x >= state[S_CHUNKS]) return;
    __shared__ uint32_t total;
    const Chunk ch = chunks[blockIdx.x];
    const TaskSplit s = splits[ch.task];
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    uint32_t mine = 0;
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += BUILD_BLOCK) mine += goes_left(s, i, order[i], p);
    atomicAdd(&total, mine);
    __syncthreads();
    if (threadIdx.x == 0) lefts[blockIdx.x] = total;
}

// Each chunk's first left and right destinations: per task, a running sum
// over its chunks in range order (one lane per task), checked against the
// split's left count.
__global__ __launch_bounds__(BUILD_BLOCK) void k_level_bases(const uint32_t* __restrict__ chunk_first, uint32_t* state,
                                                             const Chunk* __restrict__ chunks,
                                                             const TaskSplit* __restrict__ splits,
                                                             const uint32_t* __restrict__ lefts, uint32_t* lbase,
                                                             uint32_t* rbase) {
    const uint32_t nt = state[S_BIG];
    for (uint32_t t = blockIdx.x * BUILD_BLOCK + threadIdx.x; t < nt; t += gridDim.x * BUILD_BLOCK) {
        const TaskSplit s = splits[t];
        uint32_t l = 0, r = 0;
        for (uint32_t c = chunk_first[t]; c < chunk_first[t + 1]; ++c) {
            lbase[c] = s.begin + l;
            rbase[c] = s.begin + s.left + r;
            l += lefts[c];
            r += (chunks[c].end - chunks[c].begin) - lefts[c];
        }
        if (l != s.left) atomicOr(&state[S_OVERFLOW], 8u);
    }
}

__global__ __launch_bounds__(BUILD_BLOCK) void k_level_scatter(const Chunk* __restrict__ chunks,
                                                               const uint32_t* __restrict__ state,
                                                               const TaskSplit* __restrict__ splits, PrimArrays p,
                                                               const uint32_t* __restrict__ order,
                                                               const uint32_t* __restrict__ lbase,
                                                               const uint32_t* __restrict__ rbase, uint32_t* out) {
    if (blockIdx.x >= state[S_CHUNKS]) return;
    k_chunk_scatter_body(chunks[blockIdx.x], splits, p, order, lbase[blockIdx.x], rbase[blockIdx.x], out);
}

__global__ void k_level_copy(const Chunk* __restrict__ chunks, const uint32_t* __restrict__ state,
                             const uint32_t* __restrict__ src, uint32_t* dst) {
    if (blockIdx.x >= state[S_CHUNKS]) return;
    const Chunk ch = chunks[blockIdx.x];
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += BUILD_BLOCK) dst[i] = src[i];
}

// The next level's list becomes the current one.
__global__ void k_level_advance(uint32_t* state) {
    state[S_BIG] = state[S_NEXT_BIG];
    state[S_NEXT_BIG] = 0;
}

// --- subtrees: ranges of at most SMALL_TASK triangles --------------------------

// One workgroup builds the whole subtree of one small range, depth first,
// smaller child first, with the range's order held in LDS; bounds fold by
// wave shuffles and every split is decided by the whole workgroup.
__global__ __launch_bounds__(BUILD_BLOCK) void k_small_build(const Task* __restrict__ tasks, const uint32_t* counts,
                                                             PrimArrays p, uint32_t* order, BinaryNode* nodes,
                                                             uint32_t* state) {
    __shared__ uint32_t ord[SMALL_TASK];
    __shared__ uint32_t tmp[SMALL_TASK];
    __shared__ uint32_t red[TB];
    __shared__ uint32_t wave_red[(BUILD_BLOCK / 64) * TB];
    __shared__ uint32_t lb[TBIN];
    __shared__ Box3 fbins[3][BVH_BINS];
    __shared__ uint32_t fcounts[3][BVH_BINS];
    __shared__ float red_cost[2];
    __shared__ uint32_t red_idx[2], red_left[2];
    __shared__ Decision dec;
    __shared__ uint32_t stack_b[SMALL_STACK], stack_e[SMALL_STACK], stack_n[SMALL_STACK];
    __shared__ uint32_t sp, cur_b, cur_e, cur_node, left;
    __shared__ int live_leaf;  // 1: make a leaf
    __shared__ int by_bin;     // 1: partition by bin, 0: by index
    __shared__ int axis, split;
    __shared__ float cb_lo[3], cb_scale[3];
    __shared__ int cb_live[3];
    __shared__ uint32_t scan[BUILD_BLOCK];

    if (blockIdx.x >= counts[0]) return;
    const Task t = tasks[blockIdx.x];
    const uint32_t n = t.end - t.begin;
    for (uint32_t i = threadIdx.x; i < n; i += BUILD_BLOCK) ord[i] = order[t.begin + i];
    if (threadIdx.x == 0) {
        sp = 1;
        stack_b[0] = 0;
        stack_e[0] = n;
        stack_n[0] = t.node;
    }
    __syncthreads();
    while (true) {
        if (threadIdx.x == 0) {
            if (sp == 0) {
                cur_node = 0xffffffffu;
            } else {
                --sp;
                cur_b = stack_b[sp];
                cur_e = stack_e[sp];
                cur_node = stack_n[sp];
            }
        }
        __syncthreads();
        if (cur_node == 0xffffffffu) break;
        const uint32_t b0 = cur_b, e0 = cur_e, m = e0 - b0;

        // Bounds.
        uint32_t v[TB];
        for (int k = 0; k < int(TB); ++k) v[k] = (k % 6) < 3 ? O_POS_INF : O_NEG_INF;
        for (uint32_t i = b0 + threadIdx.x; i < e0; i += BUILD_BLOCK) {
            const uint32_t id = ord[i];
            for (int k = 0; k < 3; ++k) {
                v[k] = min(v[k], f2o(p.lo[k][id]));
                v[3 + k] = max(v[3 + k], f2o(p.hi[k][id]));
                const uint32_t c = f2o(p.c[k][id]);
                v[6 + k] = min(v[6 + k], c);
                v[9 + k] = max(v[9 + k], c);
            }
        }
        for (uint32_t k = threadIdx.x; k < TBIN; k += BUILD_BLOCK) {
            const uint32_t f = k % 7;
            lb[k] = f == 0 ? 0u : (f <= 3 ? O_POS_INF : O_NEG_INF);
        }
        fold_bounds(v, red, wave_red);
        if (threadIdx.x < 3) {
            const int k = threadIdx.x;
            const float lo = o2f(red[6 + k]), hi = o2f(red[9 + k]);
            cb_lo[k] = lo;
            cb_live[k] = (m > 1) && (hi - lo > 0.0f);
            cb_scale[k] = cb_live[k] ? bvh_scale(lo, hi) : 0.0f;
        }
        __syncthreads();

        // Bins.
        for (uint32_t i = b0 + threadIdx.x; i < e0; i += BUILD_BLOCK) {
            const uint32_t id = ord[i];
            uint32_t box[6];
            for (int k = 0; k < 3; ++k) {
                box[k] = f2o(p.lo[k][id]);
                box[3 + k] = f2o(p.hi[k][id]);
            }
            for (int a = 0; a < 3; ++a) {
                if (!cb_live[a]) continue;
                const int bin = bvh_bin(p.c[a][id], cb_lo[a], cb_scale[a]);
                uint32_t* e = &lb[(a * BVH_BINS + bin) * 7];
                atomicAdd(&e[0], 1u);
                for (int k = 0; k < 3; ++k) {
                    atomicMin(&e[1 + k], box[k]);
                    atomicMax(&e[4 + k], box[3 + k]);
                }
            }
        }
        __syncthreads();
        if (threadIdx.x < 3 * BVH_BINS) {
            const uint32_t a = threadIdx.x / BVH_BINS, bin = threadIdx.x % BVH_BINS;
            const uint32_t* e = &lb[threadIdx.x * 7];
            fcounts[a][bin] = e[0];
            for (int k = 0; k < 3; ++k) {
                fbins[a][bin].lo[k] = o2f(e[1 + k]);
                fbins[a][bin].hi[k] = o2f(e[4 + k]);
            }
        }
        __syncthreads();

        // Decide, exactly as the host builder does.
        Box3 box, cbox;
        for (int k = 0; k < 3; ++k) {
            box.lo[k] = o2f(red[k]);
            box.hi[k] = o2f(red[3 + k]);
            cbox.lo[k] = o2f(red[6 + k]);
            cbox.hi[k] = o2f(red[9 + k]);
        }
        decide_block(fbins, fcounts, box, cbox, m, red_cost, red_idx, red_left, &dec);
        if (threadIdx.x == 0) {
            BinaryNode out;
            out.box = box;
            live_leaf = dec.d.leaf;
            if (dec.d.leaf) {
                out.first = t.begin + b0;
                out.count = m;
                out.left = out.right = 0;
                nodes[cur_node] = out;
            } else {
                uint32_t l = m / 2;
                int bin_mode = 0;
                if (dec.d.axis >= 0 && dec.left > 0 && dec.left < m) {
                    l = dec.left;
                    bin_mode = 1;
                }
                left = l;
                by_bin = bin_mode;
                axis = dec.d.axis;
                split = dec.d.split;
                const uint32_t ids = atomicAdd(&state[S_NODES], 2u);
                out.left = ids;
                out.right = ids + 1;
                out.first = 0;
                out.count = 0;
                nodes[cur_node] = out;
                // Smaller child on top, so it is built first.
                const uint32_t lb0 = b0, le0 = b0 + l, rb0 = b0 + l, re0 = e0;
                const bool left_smaller = l <= m - l;
                if (sp + 2 > SMALL_STACK) {
                    atomicOr(&state[S_OVERFLOW], 16u);
                } else {
                    stack_b[sp] = left_smaller ? rb0 : lb0;
                    stack_e[sp] = left_smaller ? re0 : le0;
                    stack_n[sp] = left_smaller ? ids + 1 : ids;
                    ++sp;
                    stack_b[sp] = left_smaller ? lb0 : rb0;
                    stack_e[sp] = left_smaller ? le0 : re0;
                    stack_n[sp] = left_smaller ? ids : ids + 1;
                    ++sp;
                }
            }
        }
        __syncthreads();
        if (live_leaf || !by_bin) continue;  // leaves and index splits keep the order

        // Stable partition of [b0, e0) by bin.
        const uint32_t per = (m + BUILD_BLOCK - 1) / BUILD_BLOCK;
        const uint32_t r0 = min(b0 + threadIdx.x * per, e0);
        const uint32_t r1 = min(r0 + per, e0);
        uint32_t mine = 0;
        for (uint32_t i = r0; i < r1; ++i) mine += bvh_bin(p.c[axis][ord[i]], cb_lo[axis], cb_scale[axis]) <= split;
        scan[threadIdx.x] = mine;
        __syncthreads();
        for (uint32_t off = 1; off < BUILD_BLOCK; off <<= 1) {
            const uint32_t add = threadIdx.x >= off ? scan[threadIdx.x - off] : 0u;
            __syncthreads();
            scan[threadIdx.x] += add;
            __syncthreads();
        }
        const uint32_t lefts_before = scan[threadIdx.x] - mine;
        uint32_t l = b0 + lefts_before;
        uint32_t r = b0 + left + ((r0 - b0) - lefts_before);
        for (uint32_t i = r0; i < r1; ++i) {
            const uint32_t id = ord[i];
            if (bvh_bin(p.c[axis][id], cb_lo[axis], cb_scale[axis]) <= split) tmp[l++] = id;
            else tmp[r++] = id;
        }
        __syncthreads();
        for (uint32_t i = b0 + threadIdx.x; i < e0; i += BUILD_BLOCK) ord[i] = tmp[i];
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < n; i += BUILD_BLOCK) order[t.begin + i] = ord[i];
}

// --- the collapse to wide nodes, on the device ----------------------------------
//
// The host collapse (cn_bvh.cpp) numbers wide nodes in depth-first preorder,
// children in slot order. Here the same nodes are found level by level: each
// wide node at depth d opens its binary root greedily (the internal child of
// largest surface area, until W children), giving its children fresh wide
// ids and listing them for depth d + 1; then subtree sizes run bottom-up and
// preorder indices top-down (a node's k-th internal child starts after the
// node and the subtrees of its earlier children), and every node is written
// at its index.

template <int W>
struct WideTmp {
    uint32_t kid[W];       // binary node of each child
    uint32_t kid_wide[W];  // wide id of each internal child, or CHILD_EMPTY for a leaf child
    uint32_t count, size, index, depth;
};

struct ListEntry {
    uint32_t binary, wide;
};

template <int W>
__global__ __launch_bounds__(BUILD_BLOCK) void k_collapse_level(const BinaryNode* __restrict__ nodes,
                                                                const uint32_t* __restrict__ level_bounds,
                                                                uint32_t depth, ListEntry* list, WideTmp<W>* wide,
                                                                uint32_t* state, uint32_t list_capacity) {
    const uint32_t begin = level_bounds[depth], end = level_bounds[depth + 1];
    for (uint32_t i = begin + blockIdx.x * BUILD_BLOCK + threadIdx.x; i < end; i += gridDim.x * BUILD_BLOCK) {
        const ListEntry e = list[i];
        WideTmp<W> w;
        uint32_t count = 0;
        const BinaryNode& root = nodes[e.binary];
        if (root.count > 0) {
            w.kid[count++] = e.binary;
        } else {
            w.kid[count++] = root.left;
            w.kid[count++] = root.right;
            while (count < uint32_t(W)) {
                int open = -1;
                float best = -1.0f;
                for (uint32_t k = 0; k < count; ++k) {
                    const BinaryNode& c = nodes[w.kid[k]];
                    if (c.count == 0 && c.box.area() > best) {
                        best = c.box.area();
                        open = int(k);
                    }
                }
                if (open < 0) break;
                const BinaryNode& o = nodes[w.kid[open]];
                w.kid[open] = o.left;
                w.kid[count++] = o.right;
            }
        }
        for (uint32_t k = 0; k < count; ++k) {
            w.kid_wide[k] = CHILD_EMPTY;
            if (nodes[w.kid[k]].count == 0) {
                const uint32_t id = atomicAdd(&state[S_WIDE], 1u);
                const uint32_t at = atomicAdd(&state[S_LIST], 1u);
                if (at < list_capacity) list[at] = ListEntry{w.kid[k], id};
                else atomicOr(&state[S_OVERFLOW], 32u);
                w.kid_wide[k] = id;
            }
        }
        w.count = count;
        w.size = 0;
        w.index = 0;
        w.depth = depth + 1;
        wide[e.wide] = w;
    }
}

// The list's end becomes the next depth's bound.
__global__ void k_collapse_close(uint32_t* level_bounds, uint32_t depth, const uint32_t* state) {
    level_bounds[depth + 2] = state[S_LIST];
}

template <int W>
__global__ __launch_bounds__(BUILD_BLOCK) void k_collapse_sizes(const uint32_t* __restrict__ level_bounds,
                                                                uint32_t depth, const ListEntry* __restrict__ list,
                                                                WideTmp<W>* wide) {
    const uint32_t begin = level_bounds[depth], end = level_bounds[depth + 1];
    for (uint32_t i = begin + blockIdx.x * BUILD_BLOCK + threadIdx.x; i < end; i += gridDim.x * BUILD_BLOCK) {
        WideTmp<W>& w = wide[list[i].wide];
        uint32_t size = 1;
        for (uint32_t k = 0; k < w.count; ++k)
            if (w.kid_wide[k] != CHILD_EMPTY) size += wide[w.kid_wide[k]].size;
        w.size = size;
    }
}

template <int W>
__global__ __launch_bounds__(BUILD_BLOCK) void k_collapse_index(const uint32_t* __restrict__ level_bounds,
                                                                uint32_t depth, const ListEntry* __restrict__ list,
                                                                WideTmp<W>* wide) {
    const uint32_t begin = level_bounds[depth], end = level_bounds[depth + 1];
    for (uint32_t i = begin + blockIdx.x * BUILD_BLOCK + threadIdx.x; i < end; i += gridDim.x * BUILD_BLOCK) {
        const WideTmp<W>& w = wide[list[i].wide];
        uint32_t at = w.index + 1;
        for (uint32_t k = 0; k < w.count; ++k) {
            if (w.kid_wide[k] == CHILD_EMPTY) continue;
            wide[w.kid_wide[k]].index = at;
            at += wide[w.kid_wide[k]].size;
        }
    }
}

__device__ __forceinline__ uint32_t child_ref(const BinaryNode* nodes, uint32_t binary, uint32_t kid_wide,
                                              const uint32_t* indices) {
    const BinaryNode& b = nodes[binary];
    return b.count > 0 ? leaf_encode(b.first, b.count) : indices[kid_wide];
}

// One wide node's grid origin, power-of-two step and byte planes on each
// axis (quantize_node / Collapser8::quantize_wide in cn_bvh.cpp, the same
// double arithmetic): lo and hi are the padded child boxes.
__device__ void quantize_axis(const float* lo, const float* hi, int count, float* origin, uint32_t* exps_byte,
                              uint32_t* qlo, uint32_t* qhi) {
    double a = INFINITY, b = -INFINITY;
    for (int c = 0; c < count; ++c) {
        a = fmin(a, double(lo[c]));
        b = fmax(b, double(hi[c]));
    }
    if (!(a <= b)) a = b = 0.0;
    *origin = float(a);
    const double extent = b - double(*origin);
    int e = -126;
    if (extent > 0.0) {
        frexp(extent / 255.0, &e);
        e = e > -126 ? e : -126;
        while (ldexp(255.0, e) < extent) ++e;
        while (e > -126 && ldexp(255.0, e - 1) >= extent) --e;
    }
    const double step = ldexp(1.0, e);
    for (int c = 0; c < count; ++c) {
        const uint32_t ql = uint32_t(floor((double(lo[c]) - double(*origin)) / step));
        const uint32_t qh = uint32_t(ceil((double(hi[c]) - double(*origin)) / step));
        qlo[c] = ql < 255u ? ql : 255u;
        qhi[c] = qh < 255u ? qh : 255u;
    }
    *exps_byte = uint32_t(e + 127);
}

// Every 4-wide node at its preorder index: the full-precision record (the
// host collapse's set_child with the same padding) and its quantized form.
__global__ __launch_bounds__(BUILD_BLOCK) void k_collapse_write4(const BinaryNode* __restrict__ nodes,
                                                                 const WideTmp<4>* __restrict__ wide,
                                                                 const uint32_t* __restrict__ indices,
                                                                 const uint32_t* state, const uint32_t* max_abs_bits,
                                                                 BvhNode4* full, BvhNode4Q* quant) {
    const uint32_t count = state[S_WIDE];
    const float pad = 4e-6f * fmaxf(__uint_as_float(*max_abs_bits), 1e-3f);
    for (uint32_t i = blockIdx.x * BUILD_BLOCK + threadIdx.x; i < count; i += gridDim.x * BUILD_BLOCK) {
        const WideTmp<4>& w = wide[i];
        BvhNode4 n;
        float* lx = &n.lo_x.x;
        float* hx = &n.hi_x.x;
        float* ly = &n.lo_y.x;
        float* hy = &n.hi_y.x;
        float* lz = &n.lo_z.x;
        float* hz = &n.hi_z.x;
        uint32_t* ch = &n.child.x;
        for (int k = 0; k < 4; ++k) {
            lx[k] = ly[k] = lz[k] = INFINITY;
            hx[k] = hy[k] = hz[k] = -INFINITY;
            ch[k] = CHILD_EMPTY;
        }
        for (uint32_t k = 0; k < w.count; ++k) {
            const Box3& b = nodes[w.kid[k]].box;
            lx[k] = b.lo[0] - pad;
            hx[k] = b.hi[0] + pad;
            ly[k] = b.lo[1] - pad;
            hy[k] = b.hi[1] + pad;
            lz[k] = b.lo[2] - pad;
            hz[k] = b.hi[2] + pad;
            ch[k] = child_ref(nodes, w.kid[k], w.kid_wide[k], indices);
        }
        n.meta = make_uint4(w.depth, 0, 0, 0);
        full[w.index] = n;
        BvhNode4Q q{};
        const float* los[3] = {lx, ly, lz};
        const float* his[3] = {hx, hy, hz};
        float origin[3], steps[3];
        uint32_t planes_lo[3] = {0, 0, 0}, planes_hi[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a) {
            uint32_t ql[4] = {0, 0, 0, 0}, qh[4] = {0, 0, 0, 0}, eb;
            quantize_axis(los[a], his[a], int(w.count), &origin[a], &eb, ql, qh);
            for (int c = 0; c < 4; ++c) {
                planes_lo[a] |= ql[c] << (8 * c);
                planes_hi[a] |= qh[c] << (8 * c);
            }
            steps[a] = __uint_as_float(eb << 23);  // the biased exponent's power of two
        }
        q.ox = origin[0];
        q.oy = origin[1];
        q.oz = origin[2];
        q.step_x = steps[0];
        q.step_y = steps[1];
        q.step_z = steps[2];
        q.lo_x = planes_lo[0];
        q.hi_x = planes_hi[0];
        q.lo_y = planes_lo[1];
        q.hi_y = planes_hi[1];
        q.lo_z = planes_lo[2];
        q.hi_z = planes_hi[2];
        q.child = n.child;
        quant[w.index] = q;
    }
}

__global__ __launch_bounds__(BUILD_BLOCK) void k_collapse_write8(const BinaryNode* __restrict__ nodes,
                                                                 const WideTmp<8>* __restrict__ wide,
                                                                 const uint32_t* __restrict__ indices,
                                                                 const uint32_t* state, const uint32_t* max_abs_bits,
                                                                 BvhNode8Q* quant) {
    const uint32_t count = state[S_WIDE];
    const float pad = 4e-6f * fmaxf(__uint_as_float(*max_abs_bits), 1e-3f);
    for (uint32_t i = blockIdx.x * BUILD_BLOCK + threadIdx.x; i < count; i += gridDim.x * BUILD_BLOCK) {
        const WideTmp<8>& w = wide[i];
        float lo[3][8], hi[3][8];
        BvhNode8Q q{};
        for (uint32_t k = 0; k < w.count; ++k) {
            const Box3& b = nodes[w.kid[k]].box;
            for (int a = 0; a < 3; ++a) {
                lo[a][k] = b.lo[a] - pad;
                hi[a][k] = b.hi[a] + pad;
            }
        }
        uint32_t* planes_lo[3] = {q.lo_x, q.lo_y, q.lo_z};
        uint32_t* planes_hi[3] = {q.hi_x, q.hi_y, q.hi_z};
        float origin[3];
        for (int a = 0; a < 3; ++a) {
            uint32_t ql[8] = {}, qh[8] = {}, eb;
            quantize_axis(lo[a], hi[a], int(w.count), &origin[a], &eb, ql, qh);
            for (uint32_t c = 0; c < w.count; ++c) {
                planes_lo[a][c / 4] |= ql[c] << (8 * (c % 4));
                planes_hi[a][c / 4] |= qh[c] << (8 * (c % 4));
            }
            q.exps |= eb << (8 * a);
        }
        q.ox = origin[0];
        q.oy = origin[1];
        q.oz = origin[2];
        for (uint32_t c = 0; c < 8; ++c)
            q.child[c] = c < w.count ? child_ref(nodes, w.kid[c], w.kid_wide[c], indices) : CHILD_EMPTY;
        q.depth = w.depth;
        quant[w.index] = q;
    }
}

// Preorder index per wide id, for the children's refs.
template <int W>
__global__ __launch_bounds__(BUILD_BLOCK) void k_collapse_indices(const WideTmp<W>* __restrict__ wide,
                                                                  const uint32_t* state, uint32_t* indices) {
    const uint32_t count = state[S_WIDE];
    for (uint32_t i = blockIdx.x * BUILD_BLOCK + threadIdx.x; i < count; i += gridDim.x * BUILD_BLOCK)
        indices[i] = wide[i].index;
}

__global__ __launch_bounds__(BUILD_BLOCK) void k_leaf_triangles(const WorldTriangle* __restrict__ in,
                                                                const uint32_t* __restrict__ order, uint32_t n,
                                                                BvhTriangle* out) {
    const uint32_t i = blockIdx.x * BUILD_BLOCK + threadIdx.x;
    if (i >= n) return;
    const WorldTriangle t = in[order[i]];
    BvhTriangle r;
    r.v0 = make_float4(t.v0.x, t.v0.y, t.v0.z, __uint_as_float(t.instance));
    r.v1 = make_float4(t.v1.x, t.v1.y, t.v1.z, __uint_as_float(t.primitive));
    r.v2 = make_float4(t.v2.x, t.v2.y, t.v2.z, __uint_as_float(t.flags));
    out[i] = r;
}

// Device allocations freed on every exit path.
struct Scratch {
    std::vector<void*> ptrs;
    ~Scratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    T* get(size_t n) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
};

uint32_t blocks_for(size_t n) { return uint32_t((n + BUILD_BLOCK - 1) / BUILD_BLOCK); }

}  // namespace

#define BUILD_CHECK(call)                                                                 \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            *error = std::string("device BVH build: ") + #call + ": " + hipGetErrorString(e_); \
            return false;                                                                 \
        }                                                                                 \
    } while (0)

namespace {

// The collapse of the binary tree to W-wide nodes, on the device. Fills
// `*depth` (levels of the wide tree) and `*count`; returns false with
// *too_deep set when the tree has more than `max_levels` levels.
template <int W>
bool collapse_device(const BinaryNode* d_nodes, uint32_t n, uint32_t* state, uint32_t max_levels, Scratch& s,
                     WideTmp<W>** wide_out, uint32_t** indices_out, uint32_t* depth, uint32_t* count, bool* too_deep,
                     std::string* error) {
    const uint32_t capacity = std::max<uint32_t>(n, 1);
    WideTmp<W>* wide = s.get<WideTmp<W>>(capacity);
    ListEntry* list = s.get<ListEntry>(capacity);
    uint32_t* bounds = s.get<uint32_t>(max_levels + 3);
    uint32_t* indices = s.get<uint32_t>(capacity);
    if (!wide || !list || !bounds || !indices) {
        *error = "device BVH build: out of device memory";
        return false;
    }
    const ListEntry root{0, 0};
    const uint32_t init_bounds[2] = {0, 1};
    const uint32_t init_state[2] = {1, 1};  // S_WIDE, S_LIST
    BUILD_CHECK(hipMemcpy(list, &root, sizeof(root), hipMemcpyHostToDevice));
    BUILD_CHECK(hipMemset(bounds, 0, (max_levels + 3) * sizeof(uint32_t)));
    BUILD_CHECK(hipMemcpy(bounds, init_bounds, sizeof(init_bounds), hipMemcpyHostToDevice));
    BUILD_CHECK(hipMemcpy(state + S_WIDE, init_state, sizeof(init_state), hipMemcpyHostToDevice));
    const uint32_t grid = std::min<uint32_t>(blocks_for(capacity), 2048);
    for (uint32_t d = 0; d <= max_levels; ++d) {
        hipLaunchKernelGGL(k_collapse_level<W>, dim3(grid), dim3(BUILD_BLOCK), 0, nullptr, d_nodes, bounds, d, list,
                           wide, state, capacity);
        hipLaunchKernelGGL(k_collapse_close, dim3(1), dim3(1), 0, nullptr, bounds, d, state);
    }
    BUILD_CHECK(hipGetLastError());
    std::vector<uint32_t> b(max_levels + 3);
    BUILD_CHECK(hipMemcpy(b.data(), bounds, b.size() * 4, hipMemcpyDeviceToHost));
    uint32_t st[S_WORDS];
    BUILD_CHECK(hipMemcpy(st, state, sizeof(st), hipMemcpyDeviceToHost));
    if (st[S_OVERFLOW] & 32u) {
        *error = "device BVH build: collapse list overflow";
        return false;
    }
    uint32_t levels = 0;
    while (levels <= max_levels && b[levels + 1] > b[levels]) ++levels;
    *too_deep = levels > max_levels;
    *depth = levels;
    *count = st[S_WIDE];
    if (*too_deep) return false;
    for (uint32_t d = levels; d-- > 0;)
        hipLaunchKernelGGL(k_collapse_sizes<W>, dim3(grid), dim3(BUILD_BLOCK), 0, nullptr, bounds, d, list, wide);
    for (uint32_t d = 0; d < levels; ++d)
        hipLaunchKernelGGL(k_collapse_index<W>, dim3(grid), dim3(BUILD_BLOCK), 0, nullptr, bounds, d, list, wide);
    hipLaunchKernelGGL(k_collapse_indices<W>, dim3(grid), dim3(BUILD_BLOCK), 0, nullptr, wide, state, indices);
    BUILD_CHECK(hipGetLastError());
    *wide_out = wide;
    *indices_out = indices;
    return true;
}

}  // namespace

bool build_bvh_device(const std::vector<WorldTriangle>& input, uint32_t max_stack, Bvh4* out, DeviceBvh* dev,
                      std::string* error) {
    const auto t0 = std::chrono::steady_clock::now();
    *dev = DeviceBvh{};
    // The outputs are freed again unless the build finishes.
    struct Outputs {
        DeviceBvh* d;
        bool keep = false;
        ~Outputs() {
            if (keep) return;
            for (void* p : {static_cast<void*>(d->tris), static_cast<void*>(d->nodes), static_cast<void*>(d->nodes8)})
                if (p) (void)hipFree(p);
            *d = DeviceBvh{};
        }
    } outputs{dev};
    const uint32_t n = uint32_t(input.size());
    if (n == 0) {
        *error = "a scene needs at least one triangle";
        return false;
    }
    if (n > LEAF_FIRST_MASK) {
        *error = "too many triangles for the BVH leaf encoding";
        return false;
    }
    static_assert(sizeof(WorldTriangle) == 48, "uploaded as is");
    Scratch s;
    WorldTriangle* d_in = s.get<WorldTriangle>(n);
    PrimArrays p;
    for (int k = 0; k < 3; ++k) {
        p.lo[k] = s.get<float>(n);
        p.hi[k] = s.get<float>(n);
        p.c[k] = s.get<float>(n);
    }
    uint32_t* order = s.get<uint32_t>(n);
    uint32_t* order_tmp = s.get<uint32_t>(n);
    uint32_t* state = s.get<uint32_t>(S_WORDS);
    BinaryNode* d_nodes = s.get<BinaryNode>(2 * size_t(n));
    // Task lists: big tasks of one level are disjoint ranges of more than
    // SMALL_TASK triangles; small tasks are disjoint ranges.
    const uint32_t big_cap = n / SMALL_TASK + 2;
    const uint32_t small_cap = n + 1;
    const uint32_t chunk_cap = n / CHUNK + big_cap + 1;
    Task* big[2] = {s.get<Task>(big_cap), s.get<Task>(big_cap)};
    Task* small = s.get<Task>(small_cap);
    Chunk* chunks = s.get<Chunk>(chunk_cap);
    uint32_t* chunk_first = s.get<uint32_t>(big_cap + 1);
    uint32_t* tb = s.get<uint32_t>(size_t(big_cap) * TB);
    uint32_t* bins = s.get<uint32_t>(size_t(big_cap) * TBIN);
    TaskSplit* splits = s.get<TaskSplit>(big_cap);
    uint32_t* lefts = s.get<uint32_t>(chunk_cap);
    uint32_t* lbase = s.get<uint32_t>(chunk_cap);
    uint32_t* rbase = s.get<uint32_t>(chunk_cap);
    bool ok = d_in && order && order_tmp && state && d_nodes && big[0] && big[1] && small && chunks && chunk_first &&
              tb && bins && splits && lefts && lbase && rbase;
    for (int k = 0; k < 3; ++k) ok = ok && p.lo[k] && p.hi[k] && p.c[k];
    if (!ok) {
        *error = "device BVH build: out of device memory";
        return false;
    }
    BUILD_CHECK(hipMemcpy(d_in, input.data(), sizeof(WorldTriangle) * n, hipMemcpyHostToDevice));
    {
        uint32_t init[S_WORDS] = {};
        init[S_NODES] = 1;
        const Task root{0, n, 0, 0};
        if (n > SMALL_TASK) {
            init[S_BIG] = 1;
            BUILD_CHECK(hipMemcpy(big[0], &root, sizeof(root), hipMemcpyHostToDevice));
        } else {
            init[S_SMALL] = 1;
            BUILD_CHECK(hipMemcpy(small, &root, sizeof(root), hipMemcpyHostToDevice));
        }
        BUILD_CHECK(hipMemcpy(state, init, sizeof(init), hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(k_prims, dim3(blocks_for(n)), dim3(BUILD_BLOCK), 0, nullptr, d_in, n, p, order,
                       state + S_MAX_ABS, state + S_BAD);
    BUILD_CHECK(hipGetLastError());

    // Top levels, eight at a time between reads of the device's task count:
    // every decision is made on the device from device-resident bins.
    uint32_t st[S_WORDS];
    const uint32_t init_grid = std::min<uint32_t>(blocks_for(size_t(big_cap) * TBIN), 4096);
    int cur = 0;
    for (uint32_t batch = 0;; ++batch) {
        BUILD_CHECK(hipMemcpy(st, state, sizeof(st), hipMemcpyDeviceToHost));
        if (st[S_BAD]) {
            *error = "scene geometry holds a non-finite vertex";
            return false;
        }
        if (st[S_OVERFLOW]) {
            *error = "device BVH build: task list overflow (" + std::to_string(st[S_OVERFLOW]) + ")";
            return false;
        }
        if (st[S_BIG] == 0) break;
        if (batch > n) {
            *error = "device BVH build: the top levels did not finish";
            return false;
        }
        for (int level = 0; level < 8; ++level) {
            hipLaunchKernelGGL(k_level_chunks, dim3(1), dim3(BUILD_BLOCK), 0, nullptr, big[cur], state, chunks,
                               chunk_first, chunk_cap);
            hipLaunchKernelGGL(k_level_init, dim3(init_grid), dim3(BUILD_BLOCK), 0, nullptr, state, tb, bins);
            hipLaunchKernelGGL(k_level_bounds, dim3(chunk_cap), dim3(BUILD_BLOCK), 0, nullptr, chunks, state, p,
                               order, tb);
            hipLaunchKernelGGL(k_level_bins, dim3(chunk_cap), dim3(BUILD_BLOCK), 0, nullptr, chunks, state, p, order,
                               tb, bins);
            hipLaunchKernelGGL(k_level_decide, dim3(big_cap), dim3(BUILD_BLOCK), 0, nullptr, big[cur], state, tb,
                               bins, splits, d_nodes, big[1 - cur], small, big_cap, small_cap);
            hipLaunchKernelGGL(k_level_count, dim3(chunk_cap), dim3(BUILD_BLOCK), 0, nullptr, chunks, state, splits,
                               p, order, lefts);
            hipLaunchKernelGGL(k_level_bases, dim3(blocks_for(big_cap)), dim3(BUILD_BLOCK), 0, nullptr, chunk_first,
                               state, chunks, splits, lefts, lbase, rbase);
            hipLaunchKernelGGL(k_level_scatter, dim3(chunk_cap), dim3(BUILD_BLOCK), 0, nullptr, chunks, state, splits,
                               p, order, lbase, rbase, order_tmp);
            hipLaunchKernelGGL(k_level_copy, dim3(chunk_cap), dim3(BUILD_BLOCK), 0, nullptr, chunks, state,
                               order_tmp, order);
            hipLaunchKernelGGL(k_level_advance, dim3(1), dim3(1), 0, nullptr, state);
            cur = 1 - cur;
        }
        BUILD_CHECK(hipGetLastError());
    }

    // Subtrees, one workgroup each.
    if (st[S_SMALL] > 0) {
        hipLaunchKernelGGL(k_small_build, dim3(st[S_SMALL]), dim3(BUILD_BLOCK), 0, nullptr, small, state + S_SMALL, p,
                           order, d_nodes, state);
        BUILD_CHECK(hipGetLastError());
    }
    BUILD_CHECK(hipMemcpy(st, state, sizeof(st), hipMemcpyDeviceToHost));
    if (st[S_OVERFLOW]) {
        *error = "device BVH build: subtree stack overflow";
        return false;
    }
    if (st[S_NODES] > 2 * size_t(n)) {
        *error = "device BVH build: node count out of range";
        return false;
    }
    float max_abs;
    std::memcpy(&max_abs, &st[S_MAX_ABS], 4);
    (void)max_abs;

    // Collapse to 4-wide nodes (preorder numbering, as the host collapse),
    // full precision and quantized, then to 8-wide quantized nodes.
    WideTmp<4>* wide4 = nullptr;
    uint32_t* idx4 = nullptr;
    uint32_t depth4 = 0, count4 = 0;
    bool deep = false;
    if (!collapse_device<4>(d_nodes, n, state, max_stack / 3, s, &wide4, &idx4, &depth4, &count4, &deep, error)) {
        if (deep) *error = "BVH too deep for the traversal stack (depth " + std::to_string(depth4) + ")";
        return false;
    }
    BvhNode4* full4 = s.get<BvhNode4>(count4);
    BvhNode4Q* quant4 = nullptr;
    BUILD_CHECK(hipMalloc(reinterpret_cast<void**>(&quant4), sizeof(BvhNode4Q) * count4));
    dev->nodes = quant4;
    if (!full4) {
        *error = "device BVH build: out of device memory";
        return false;
    }
    const uint32_t grid4 = std::min<uint32_t>(blocks_for(count4), 2048);
    hipLaunchKernelGGL(k_collapse_write4, dim3(grid4), dim3(BUILD_BLOCK), 0, nullptr, d_nodes, wide4, idx4, state,
                       state + S_MAX_ABS, full4, quant4);
    BUILD_CHECK(hipGetLastError());
    out->nodes.resize(count4);
    BUILD_CHECK(hipMemcpy(out->nodes.data(), full4, sizeof(BvhNode4) * count4, hipMemcpyDeviceToHost));
    out->max_depth = depth4;

    WideTmp<8>* wide8 = nullptr;
    uint32_t* idx8 = nullptr;
    uint32_t depth8 = 0, count8 = 0;
    out->nodes8.clear();
    out->max_depth8 = 0;
    if (collapse_device<8>(d_nodes, n, state, max_stack / 7, s, &wide8, &idx8, &depth8, &count8, &deep, error)) {
        BvhNode8Q* quant8 = nullptr;
        BUILD_CHECK(hipMalloc(reinterpret_cast<void**>(&quant8), sizeof(BvhNode8Q) * count8));
        dev->nodes8 = quant8;
        const uint32_t grid8 = std::min<uint32_t>(blocks_for(count8), 2048);
        hipLaunchKernelGGL(k_collapse_write8, dim3(grid8), dim3(BUILD_BLOCK), 0, nullptr, d_nodes, wide8, idx8, state,
                           state + S_MAX_ABS, quant8);
        BUILD_CHECK(hipGetLastError());
        out->nodes8.resize(count8);
        BUILD_CHECK(hipMemcpy(out->nodes8.data(), quant8, sizeof(BvhNode8Q) * count8, hipMemcpyDeviceToHost));
        out->max_depth8 = depth8;
    } else if (!deep) {
        return false;
    }

    BvhTriangle* d_tris = nullptr;
    BUILD_CHECK(hipMalloc(reinterpret_cast<void**>(&d_tris), sizeof(BvhTriangle) * n));
    dev->tris = d_tris;
    hipLaunchKernelGGL(k_leaf_triangles, dim3(blocks_for(n)), dim3(BUILD_BLOCK), 0, nullptr, d_in, order, n, d_tris);
    BUILD_CHECK(hipGetLastError());
    BUILD_CHECK(hipDeviceSynchronize());
    out->tris.clear();
    out->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    outputs.keep = true;
    return true;
}

namespace {

// One thread per leaf-order triangle: the flags word of its instance.
__global__ __launch_bounds__(BUILD_BLOCK) void k_patch_flags(BvhTriangle* __restrict__ tris, uint32_t n,
                                                             const uint32_t* __restrict__ instance_flags,
                                                             uint32_t instances) {
    const uint32_t i = blockIdx.x * BUILD_BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t inst = __float_as_uint(tris[i].v0.w);
    if (inst < instances) tris[i].v2.w = __uint_as_float(instance_flags[inst]);
}

// One thread per node: each plane byte q of the quantized node becomes
// q * step (exact), the origin and children carried over. An empty slot's
// box is inverted to +inf .. -inf on every axis, so a walk that takes each
// axis's near plane by the direction's sign never enters it.
__global__ __launch_bounds__(BUILD_BLOCK) void k_packet_nodes(const BvhNode4Q* __restrict__ in, uint32_t n,
                                                              BvhNode4P* __restrict__ out) {
    const uint32_t i = blockIdx.x * BUILD_BLOCK + threadIdx.x;
    if (i >= n) return;
    const BvhNode4Q q = in[i];
    BvhNode4P p;
    p.ox = q.ox;
    p.oy = q.oy;
    p.oz = q.oz;
    p.pad = 0.0f;
    const uint32_t child[4] = {q.child.x, q.child.y, q.child.z, q.child.w};
    const float inf = __int_as_float(0x7f800000);
    for (int c = 0; c < 4; ++c) {
        const int sh = 8 * c;
        const bool empty = child[c] == CHILD_EMPTY;
        p.lo_x[c] = empty ? inf : float((q.lo_x >> sh) & 0xffu) * q.step_x;
        p.hi_x[c] = empty ? -inf : float((q.hi_x >> sh) & 0xffu) * q.step_x;
        p.lo_y[c] = empty ? inf : float((q.lo_y >> sh) & 0xffu) * q.step_y;
        p.hi_y[c] = empty ? -inf : float((q.hi_y >> sh) & 0xffu) * q.step_y;
        p.lo_z[c] = empty ? inf : float((q.lo_z >> sh) & 0xffu) * q.step_z;
        p.hi_z[c] = empty ? -inf : float((q.hi_z >> sh) & 0xffu) * q.step_z;
        p.child[c] = child[c];
    }
    out[i] = p;
}

}  // namespace

bool decode_packet_nodes(const BvhNode4Q* nodes, uint32_t count, BvhNode4P* out, std::string* error) {
    if (count == 0) return true;
    hipLaunchKernelGGL(k_packet_nodes, dim3(blocks_for(count)), dim3(BUILD_BLOCK), 0, nullptr, nodes, count, out);
    BUILD_CHECK(hipGetLastError());
    BUILD_CHECK(hipDeviceSynchronize());
    return true;
}

bool patch_triangle_flags(BvhTriangle* tris, uint32_t count, const uint32_t* instance_flags, uint32_t instances,
                          std::string* error) {
    if (count == 0) return true;
    Scratch s;
    uint32_t* d_flags = s.get<uint32_t>(instances);
    if (!d_flags) {
        *error = "patch_triangle_flags: out of device memory";
        return false;
    }
    BUILD_CHECK(hipMemcpy(d_flags, instance_flags, sizeof(uint32_t) * instances, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_patch_flags, dim3(blocks_for(count)), dim3(BUILD_BLOCK), 0, nullptr, tris, count, d_flags,
                       instances);
    BUILD_CHECK(hipGetLastError());
    BUILD_CHECK(hipDeviceSynchronize());
    return true;
}

}  // namespace cn
