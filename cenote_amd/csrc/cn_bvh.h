// cn_bvh.h — the 4-wide BVH over world-space triangles that stands in for
// the reference's driver-built TLAS/BLAS pair (gpu/accel.rs:82-163).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "cn_math.h"
#include "cn_records.h"

namespace cn {

struct WorldTriangle {
    v3 v0, v1, v2;
    uint32_t instance, primitive, flags;
};

struct Bvh4 {
    std::vector<BvhNode4> nodes;  // node 0 is the root and always internal
    std::vector<BvhTriangle> tris;  // leaf order
    uint32_t max_depth = 0;
    double build_seconds = 0.0;
};

// Binned-SAH binary build (32 bins on the three axes, leaves of <= 4
// triangles), collapsed to 4-wide nodes by repeatedly opening the child of
// largest surface area. Returns false with a message when the tree would
// be deeper than the traversal stack allows.
bool build_bvh4(const std::vector<WorldTriangle>& input, uint32_t max_stack, Bvh4* out, std::string* error);

}  // namespace cn
