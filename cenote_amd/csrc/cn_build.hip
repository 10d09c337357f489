// cn_build.hip — the BVH built on the device (the reference builds its
// acceleration structures with the driver, gpu/accel.rs:167-269).
//
// The binary tree is the host builder's (cn_bvh.cpp) bit for bit: every
// split is made by the shared bvh_decide from bins that are exact in any
// order (min/max folds, integer counts), and ranges are split by a stable
// partition. Two phases:
//
//   * top levels, ranges of more than SMALL_TASK triangles: device-wide
//     passes over 8192-triangle chunks — bounds, then 32-bin histograms on
//     three axes — the decisions on the host from the read-back bins, then a
//     stable scatter (per-chunk counts, host prefix sums, in-chunk scans);
//   * every remaining range: one 256-lane workgroup builds its whole subtree
//     in LDS, depth first, smaller child first (so its stack stays short).
//
// The binary nodes are collapsed to 4-wide nodes on the host, by the same
// collapse as the host builder; the leaf-ordered triangle records are
// written on the device.

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

__global__ void k_fill(uint32_t* a, uint32_t n, uint32_t lo_value, uint32_t hi_value, uint32_t period, uint32_t split) {
    // Element i gets lo_value when (i % period) < split, else hi_value.
    const uint32_t i = blockIdx.x * BUILD_BLOCK + threadIdx.x;
    if (i < n) a[i] = (i % period) < split ? lo_value : hi_value;
}

// Bounds of each chunk's triangles, folded into its task's bounds.
__global__ __launch_bounds__(BUILD_BLOCK) void k_chunk_bounds(const Chunk* __restrict__ chunks, PrimArrays p,
                                                              const uint32_t* __restrict__ order, uint32_t* tb) {
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
        if ((k % 6) < 3) atomicMin(&red[k], v[k]);
        else atomicMax(&red[k], v[k]);
    }
    __syncthreads();
    if (threadIdx.x < TB) {
        uint32_t* dst = tb + size_t(ch.task) * TB + threadIdx.x;
        if ((threadIdx.x % 6) < 3) atomicMin(dst, red[threadIdx.x]);
        else atomicMax(dst, red[threadIdx.x]);
    }
}

// 32-bin histograms on the three axes of each chunk, folded into its task's.
__global__ __launch_bounds__(BUILD_BLOCK) void k_chunk_bins(const Chunk* __restrict__ chunks, PrimArrays p,
                                                            const uint32_t* __restrict__ order,
                                                            const uint32_t* __restrict__ tb, uint32_t* bins) {
    __shared__ uint32_t lb[TBIN];
    const Chunk ch = chunks[blockIdx.x];
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

// Left-going triangles per chunk.
__global__ __launch_bounds__(BUILD_BLOCK) void k_chunk_count(const Chunk* __restrict__ chunks,
                                                             const TaskSplit* __restrict__ splits, PrimArrays p,
                                                             const uint32_t* __restrict__ order, uint32_t* lefts) {
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

// Stable scatter of each chunk: lanes take contiguous runs, an exclusive
// scan of the runs' left counts orders them.
__global__ __launch_bounds__(BUILD_BLOCK) void k_chunk_scatter(const Chunk* __restrict__ chunks,
                                                               const TaskSplit* __restrict__ splits, PrimArrays p,
                                                               const uint32_t* __restrict__ order,
                                                               const uint32_t* __restrict__ left_base,
                                                               const uint32_t* __restrict__ right_base,
                                                               uint32_t* out) {
    __shared__ uint32_t scan[BUILD_BLOCK];
    const Chunk ch = chunks[blockIdx.x];
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
    uint32_t lefts_before = scan[threadIdx.x] - mine;
    uint32_t rights_before = (r0 - ch.begin) - lefts_before;
    uint32_t l = left_base[blockIdx.x] + lefts_before;
    uint32_t r = right_base[blockIdx.x] + rights_before;
    for (uint32_t i = r0; i < r1; ++i) {
        const uint32_t id = order[i];
        if (goes_left(s, i, id, p)) out[l++] = id;
        else out[r++] = id;
    }
}

__global__ void k_chunk_copy(const Chunk* __restrict__ chunks, const uint32_t* __restrict__ src, uint32_t* dst) {
    const Chunk ch = chunks[blockIdx.x];
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += BUILD_BLOCK) dst[i] = src[i];
}

struct SmallTask {
    uint32_t begin, end, node, pad;
};

// One workgroup builds the whole subtree of one range of <= SMALL_TASK
// triangles, depth first, with the range's order held in LDS.
__global__ __launch_bounds__(BUILD_BLOCK) void k_small_build(const SmallTask* __restrict__ tasks, PrimArrays p,
                                                             uint32_t* order, BinaryNode* nodes,
                                                             uint32_t* node_counter, uint32_t* overflow) {
    __shared__ uint32_t ord[SMALL_TASK];
    __shared__ uint32_t tmp[SMALL_TASK];
    __shared__ uint32_t red[TB];
    __shared__ uint32_t lb[TBIN];
    __shared__ Box3 fbins[3][BVH_BINS];
    __shared__ uint32_t fcounts[3][BVH_BINS];
    __shared__ uint32_t stack_b[SMALL_STACK], stack_e[SMALL_STACK], stack_n[SMALL_STACK];
    __shared__ uint32_t sp, cur_b, cur_e, cur_node, left;
    __shared__ int live_leaf;  // 1: make a leaf
    __shared__ int by_bin;     // 1: partition by bin, 0: by index
    __shared__ int axis, split;
    __shared__ float cb_lo[3], cb_scale[3];
    __shared__ int cb_live[3];
    __shared__ uint32_t scan[BUILD_BLOCK];

    const SmallTask t = tasks[blockIdx.x];
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
        if (threadIdx.x < TB) red[threadIdx.x] = (threadIdx.x % 6) < 3 ? O_POS_INF : O_NEG_INF;
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
        for (int k = 0; k < int(TB); ++k) {
            if ((k % 6) < 3) atomicMin(&red[k], v[k]);
            else atomicMax(&red[k], v[k]);
        }
        for (uint32_t k = threadIdx.x; k < TBIN; k += BUILD_BLOCK) {
            const uint32_t f = k % 7;
            lb[k] = f == 0 ? 0u : (f <= 3 ? O_POS_INF : O_NEG_INF);
        }
        __syncthreads();
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

        // Decide (one lane), exactly as the host builder does.
        if (threadIdx.x == 0) {
            Box3 box, cbox;
            for (int k = 0; k < 3; ++k) {
                box.lo[k] = o2f(red[k]);
                box.hi[k] = o2f(red[3 + k]);
                cbox.lo[k] = o2f(red[6 + k]);
                cbox.hi[k] = o2f(red[9 + k]);
            }
            const SplitDecision d = bvh_decide(box, cbox, fbins, fcounts, m);
            BinaryNode out;
            out.box = box;
            live_leaf = d.leaf;
            if (d.leaf) {
                out.first = t.begin + b0;
                out.count = m;
                out.left = out.right = 0;
                nodes[cur_node] = out;
            } else {
                uint32_t l = m / 2;
                int bin_mode = 0;
                if (d.axis >= 0) {
                    const uint32_t lc = bvh_left_count(d, fcounts);
                    if (lc > 0 && lc < m) {
                        l = lc;
                        bin_mode = 1;
                    }
                }
                left = l;
                by_bin = bin_mode;
                axis = d.axis;
                split = d.split;
                const uint32_t ids = atomicAdd(node_counter, 2u);
                out.left = ids;
                out.right = ids + 1;
                out.first = 0;
                out.count = 0;
                nodes[cur_node] = out;
                // Smaller child on top, so it is built first.
                const uint32_t lb0 = b0, le0 = b0 + l, rb0 = b0 + l, re0 = e0;
                const bool left_smaller = l <= m - l;
                if (sp + 2 > SMALL_STACK) {
                    atomicOr(overflow, 1u);
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

bool build_bvh4_device(const std::vector<WorldTriangle>& input, uint32_t max_stack, Bvh4* out, BvhTriangle** tris,
                       std::string* error) {
    const auto t0 = std::chrono::steady_clock::now();
    *tris = nullptr;
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
    uint32_t* flags = s.get<uint32_t>(4);  // max |coordinate| bits, non-finite flag, node counter, stack overflow
    BinaryNode* d_nodes = s.get<BinaryNode>(2 * size_t(n));
    bool ok = d_in && order && order_tmp && flags && d_nodes;
    for (int k = 0; k < 3; ++k) ok = ok && p.lo[k] && p.hi[k] && p.c[k];
    if (!ok) {
        *error = "device BVH build: out of device memory";
        return false;
    }
    BUILD_CHECK(hipMemcpy(d_in, input.data(), sizeof(WorldTriangle) * n, hipMemcpyHostToDevice));
    BUILD_CHECK(hipMemset(flags, 0, 4 * sizeof(uint32_t)));
    hipLaunchKernelGGL(k_prims, dim3(blocks_for(n)), dim3(BUILD_BLOCK), 0, nullptr, d_in, n, p, order, flags,
                       flags + 1);
    BUILD_CHECK(hipGetLastError());
    uint32_t head[2];
    BUILD_CHECK(hipMemcpy(head, flags, 8, hipMemcpyDeviceToHost));
    if (head[1]) {
        *error = "scene geometry holds a non-finite vertex";
        return false;
    }
    float max_abs;
    std::memcpy(&max_abs, &head[0], 4);

    // Top levels.
    struct Range {
        uint32_t begin, end, node;
    };
    std::vector<Range> large, small;
    std::vector<BinaryNode> top;  // nodes 0 .. top.size() - 1, written by the host
    top.emplace_back();
    uint32_t next_node = 1;
    if (n > SMALL_TASK) large.push_back({0, n, 0});
    else small.push_back({0, n, 0});
    while (!large.empty()) {
        std::vector<Chunk> chunks;
        for (uint32_t k = 0; k < large.size(); ++k)
            for (uint32_t b = large[k].begin; b < large[k].end; b += CHUNK)
                chunks.push_back({k, b, std::min(b + CHUNK, large[k].end), 0});
        const uint32_t tasks = uint32_t(large.size()), nc = uint32_t(chunks.size());
        Scratch level;
        Chunk* d_chunks = level.get<Chunk>(nc);
        uint32_t* d_tb = level.get<uint32_t>(size_t(tasks) * TB);
        uint32_t* d_bins = level.get<uint32_t>(size_t(tasks) * TBIN);
        TaskSplit* d_splits = level.get<TaskSplit>(tasks);
        uint32_t* d_lefts = level.get<uint32_t>(nc);
        uint32_t* d_lbase = level.get<uint32_t>(nc);
        uint32_t* d_rbase = level.get<uint32_t>(nc);
        if (!d_chunks || !d_tb || !d_bins || !d_splits || !d_lefts || !d_lbase || !d_rbase) {
            *error = "device BVH build: out of device memory";
            return false;
        }
        BUILD_CHECK(hipMemcpy(d_chunks, chunks.data(), sizeof(Chunk) * nc, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_fill, dim3(blocks_for(size_t(tasks) * TB)), dim3(BUILD_BLOCK), 0, nullptr, d_tb,
                           tasks * TB, O_POS_INF, O_NEG_INF, 6u, 3u);
        hipLaunchKernelGGL(k_chunk_bounds, dim3(nc), dim3(BUILD_BLOCK), 0, nullptr, d_chunks, p, order, d_tb);
        BUILD_CHECK(hipGetLastError());
        {
            // Bin slots, 7 words each: a count from zero, then the box's lo
            // (min identity) and hi (max identity).
            std::vector<uint32_t> init(size_t(tasks) * TBIN);
            for (size_t k = 0; k < init.size(); ++k) {
                const uint32_t f = uint32_t(k % 7);
                init[k] = f == 0 ? 0u : (f <= 3 ? O_POS_INF : O_NEG_INF);
            }
            BUILD_CHECK(hipMemcpy(d_bins, init.data(), init.size() * 4, hipMemcpyHostToDevice));
        }
        hipLaunchKernelGGL(k_chunk_bins, dim3(nc), dim3(BUILD_BLOCK), 0, nullptr, d_chunks, p, order, d_tb, d_bins);
        BUILD_CHECK(hipGetLastError());
        std::vector<uint32_t> tb(size_t(tasks) * TB), bins(size_t(tasks) * TBIN);
        BUILD_CHECK(hipMemcpy(tb.data(), d_tb, tb.size() * 4, hipMemcpyDeviceToHost));
        BUILD_CHECK(hipMemcpy(bins.data(), d_bins, bins.size() * 4, hipMemcpyDeviceToHost));

        std::vector<TaskSplit> splits(tasks);
        std::vector<Range> next_large;
        for (uint32_t k = 0; k < tasks; ++k) {
            const Range& rg = large[k];
            const uint32_t m = rg.end - rg.begin;
            Box3 box, cbox;
            for (int a = 0; a < 3; ++a) {
                box.lo[a] = o2f(tb[k * TB + a]);
                box.hi[a] = o2f(tb[k * TB + 3 + a]);
                cbox.lo[a] = o2f(tb[k * TB + 6 + a]);
                cbox.hi[a] = o2f(tb[k * TB + 9 + a]);
            }
            Box3 fb[3][BVH_BINS];
            uint32_t fc[3][BVH_BINS];
            for (int a = 0; a < 3; ++a) {
                const bool live = cbox.hi[a] - cbox.lo[a] > 0.0f;
                for (int b = 0; b < BVH_BINS; ++b) {
                    const uint32_t* e = &bins[size_t(k) * TBIN + (a * BVH_BINS + b) * 7];
                    fc[a][b] = live ? e[0] : 0u;
                    for (int q = 0; q < 3; ++q) {
                        fb[a][b].lo[q] = live ? o2f(e[1 + q]) : INFINITY;
                        fb[a][b].hi[q] = live ? o2f(e[4 + q]) : -INFINITY;
                    }
                }
            }
            const SplitDecision d = bvh_decide(box, cbox, fb, fc, m);
            TaskSplit& sp = splits[k];
            sp = TaskSplit{rg.begin, m, 1u, 0u, 0, m / 2, 0.0f, 0.0f};
            if (d.axis >= 0) {
                const uint32_t lc = bvh_left_count(d, fc);
                if (lc > 0 && lc < m) {
                    sp.mode = 0;
                    sp.axis = uint32_t(d.axis);
                    sp.split = d.split;
                    sp.left = lc;
                    sp.lo = cbox.lo[d.axis];
                    sp.scale = bvh_scale(cbox.lo[d.axis], cbox.hi[d.axis]);
                }
            }
            // Ranges this large always split (a leaf holds at most BVH_LEAF_SIZE triangles).
            BinaryNode& node = top[rg.node];
            node.box = box;
            node.left = next_node;
            node.right = next_node + 1;
            node.first = node.count = 0;
            top.resize(std::max<size_t>(top.size(), next_node + 2));
            const Range kids[2] = {{rg.begin, rg.begin + sp.left, next_node},
                                   {rg.begin + sp.left, rg.end, next_node + 1}};
            next_node += 2;
            for (const Range& kid : kids) {
                if (kid.end - kid.begin > SMALL_TASK) next_large.push_back(kid);
                else small.push_back(kid);
            }
        }
        BUILD_CHECK(hipMemcpy(d_splits, splits.data(), sizeof(TaskSplit) * tasks, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_chunk_count, dim3(nc), dim3(BUILD_BLOCK), 0, nullptr, d_chunks, d_splits, p, order,
                           d_lefts);
        BUILD_CHECK(hipGetLastError());
        std::vector<uint32_t> lefts(nc), lbase(nc), rbase(nc);
        BUILD_CHECK(hipMemcpy(lefts.data(), d_lefts, nc * 4, hipMemcpyDeviceToHost));
        std::vector<uint32_t> lrun(tasks, 0), rrun(tasks, 0);
        for (uint32_t c = 0; c < nc; ++c) {
            const uint32_t k = chunks[c].task;
            lbase[c] = splits[k].begin + lrun[k];
            rbase[c] = splits[k].begin + splits[k].left + rrun[k];
            lrun[k] += lefts[c];
            rrun[k] += (chunks[c].end - chunks[c].begin) - lefts[c];
        }
        for (uint32_t k = 0; k < tasks; ++k) {
            if (lrun[k] != splits[k].left) {
                *error = "device BVH build: partition counts disagree with the bins";
                return false;
            }
        }
        BUILD_CHECK(hipMemcpy(d_lbase, lbase.data(), nc * 4, hipMemcpyHostToDevice));
        BUILD_CHECK(hipMemcpy(d_rbase, rbase.data(), nc * 4, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_chunk_scatter, dim3(nc), dim3(BUILD_BLOCK), 0, nullptr, d_chunks, d_splits, p, order,
                           d_lbase, d_rbase, order_tmp);
        hipLaunchKernelGGL(k_chunk_copy, dim3(nc), dim3(BUILD_BLOCK), 0, nullptr, d_chunks, order_tmp, order);
        BUILD_CHECK(hipGetLastError());
        BUILD_CHECK(hipDeviceSynchronize());
        large.swap(next_large);
    }
    top.resize(next_node);
    BUILD_CHECK(hipMemcpy(d_nodes, top.data(), sizeof(BinaryNode) * top.size(), hipMemcpyHostToDevice));

    // Subtrees, one workgroup each.
    if (!small.empty()) {
        std::vector<SmallTask> st(small.size());
        for (size_t k = 0; k < small.size(); ++k) st[k] = {small[k].begin, small[k].end, small[k].node, 0};
        SmallTask* d_small = s.get<SmallTask>(st.size());
        if (!d_small) {
            *error = "device BVH build: out of device memory";
            return false;
        }
        BUILD_CHECK(hipMemcpy(d_small, st.data(), sizeof(SmallTask) * st.size(), hipMemcpyHostToDevice));
        BUILD_CHECK(hipMemcpy(flags + 2, &next_node, 4, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_small_build, dim3(uint32_t(st.size())), dim3(BUILD_BLOCK), 0, nullptr, d_small, p,
                           order, d_nodes, flags + 2, flags + 3);
        BUILD_CHECK(hipGetLastError());
    }
    uint32_t tail[2];
    BUILD_CHECK(hipMemcpy(tail, flags + 2, 8, hipMemcpyDeviceToHost));
    const uint32_t node_count = small.empty() ? next_node : tail[0];
    if (tail[1]) {
        *error = "device BVH build: subtree stack overflow";
        return false;
    }
    if (node_count > 2 * size_t(n)) {
        *error = "device BVH build: node count out of range";
        return false;
    }
    std::vector<BinaryNode> bnodes(node_count);
    BUILD_CHECK(hipMemcpy(bnodes.data(), d_nodes, sizeof(BinaryNode) * node_count, hipMemcpyDeviceToHost));
    if (!collapse_to_bvh4(bnodes, max_abs, max_stack, out, error)) return false;

    BvhTriangle* d_tris = nullptr;
    BUILD_CHECK(hipMalloc(reinterpret_cast<void**>(&d_tris), sizeof(BvhTriangle) * n));
    hipLaunchKernelGGL(k_leaf_triangles, dim3(blocks_for(n)), dim3(BUILD_BLOCK), 0, nullptr, d_in, order, n, d_tris);
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipDeviceSynchronize();
    if (le != hipSuccess || se != hipSuccess) {
        (void)hipFree(d_tris);
        *error = std::string("device BVH build: ") + hipGetErrorString(le != hipSuccess ? le : se);
        return false;
    }
    *tris = d_tris;
    out->tris.clear();
    out->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

}  // namespace cn
