// cn_bvh_split.h — the binned-SAH split decision shared by the host BVH
// builder and the device builder (host orchestration and the per-workgroup
// subtree kernel). One function makes every split, so both builders grow the
// same binary tree bit for bit: bounds are min/max folds (exact in any
// order), bin counts are integers, and the SAH sweep below runs in one fixed
// operation order on either side (compiled with -ffp-contract=off).
#pragma once

#include <math.h>
#include <stdint.h>

#include <hip/hip_runtime.h>

namespace cn {

constexpr int BVH_BINS = 32;
constexpr uint32_t BVH_LEAF_SIZE = 8;
// A node step costs about two triangle tests in the traversal (four boxes,
// the child ordering and a stack operation against one watertight test):
// the SAH's traversal term, measured best of 1, 1.5 and 2 on the 2M-triangle
// benchmark scene.
constexpr float BVH_NODE_COST = 2.0f;

struct Box3 {
    float lo[3], hi[3];

    __host__ __device__ static Box3 empty() {
        Box3 b;
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = INFINITY;
            b.hi[k] = -INFINITY;
        }
        return b;
    }
    __host__ __device__ void grow(const Box3& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], b.lo[k]);
            hi[k] = fmaxf(hi[k], b.hi[k]);
        }
    }
    __host__ __device__ float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

// The bin of a centroid coordinate on an axis of positive extent.
__host__ __device__ inline int bvh_bin(float c, float lo, float scale) {
    int b = int((c - lo) * scale);
    return b < 0 ? 0 : (b > BVH_BINS - 1 ? BVH_BINS - 1 : b);
}

__host__ __device__ inline float bvh_scale(float lo, float hi) { return float(BVH_BINS) / (hi - lo); }

struct SplitDecision {
    int leaf;   // 1: make a leaf of the whole range
    int axis;   // -1: no binned split exists (split the range at its middle)
    int split;  // bins 0..split go left
};

// Binned SAH over the three axes (32 bins; leaves of <= 8 triangles when a
// leaf is no costlier than the best split, a triangle test costing 1 and a
// node step BVH_NODE_COST). `bins`/`counts` hold the
// range's per-axis bins; axes whose centroid extent is not positive are
// skipped (their bins are ignored).
__host__ __device__ inline SplitDecision bvh_decide(const Box3& box, const Box3& cbox, const Box3 (*bins)[BVH_BINS],
                                                    const uint32_t (*counts)[BVH_BINS], uint32_t n) {
    SplitDecision d{0, -1, 0};
    if (n <= 1) {
        d.leaf = 1;
        return d;
    }
    float best_cost = INFINITY;
    for (int axis = 0; axis < 3; ++axis) {
        const float extent = cbox.hi[axis] - cbox.lo[axis];
        if (!(extent > 0.0f)) continue;
        float right_area[BVH_BINS];
        uint32_t right_count[BVH_BINS];
        Box3 acc = Box3::empty();
        uint32_t cnt = 0;
        for (int b = BVH_BINS - 1; b > 0; --b) {
            acc.grow(bins[axis][b]);
            cnt += counts[axis][b];
            right_area[b] = acc.area();
            right_count[b] = cnt;
        }
        Box3 lacc = Box3::empty();
        uint32_t lcnt = 0;
        for (int b = 0; b < BVH_BINS - 1; ++b) {
            lacc.grow(bins[axis][b]);
            lcnt += counts[axis][b];
            if (lcnt == 0 || right_count[b + 1] == 0) continue;
            const float cost = lacc.area() * float(lcnt) + right_area[b + 1] * float(right_count[b + 1]);
            if (cost < best_cost) {
                best_cost = cost;
                d.axis = axis;
                d.split = b;
            }
        }
    }
    const float parent_area = box.area();
    const float leaf_cost = float(n);
    const float split_cost = BVH_NODE_COST + (parent_area > 0.0f ? best_cost / parent_area : 0.0f);
    if (n <= BVH_LEAF_SIZE && (d.axis < 0 || leaf_cost <= split_cost)) d.leaf = 1;
    return d;
}

// Primitives whose centroid falls in bins 0..split of the decision's axis,
// out of a range whose bins are `counts`.
__host__ __device__ inline uint32_t bvh_left_count(const SplitDecision& d, const uint32_t (*counts)[BVH_BINS]) {
    uint32_t left = 0;
    for (int b = 0; b <= d.split; ++b) left += counts[d.axis][b];
    return left;
}

// A triangle's bounds with negative zeros made positive (so that min/max
// folds cannot depend on operand order) and its box centroid.
__host__ __device__ inline void bvh_prim_bounds(const float v[3][3], Box3* box, float c[3]) {
    *box = Box3::empty();
    for (int q = 0; q < 3; ++q) {
        for (int k = 0; k < 3; ++k) {
            const float x = v[q][k] + 0.0f;
            box->lo[k] = fminf(box->lo[k], x);
            box->hi[k] = fmaxf(box->hi[k], x);
        }
    }
    for (int k = 0; k < 3; ++k) c[k] = 0.5f * (box->lo[k] + box->hi[k]);
}

}  // namespace cn
