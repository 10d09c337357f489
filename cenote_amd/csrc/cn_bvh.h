// cn_bvh.h — the 4-wide BVH over world-space triangles that stands in for
// the reference's driver-built TLAS/BLAS pair (gpu/accel.rs:82-163).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "cn_bvh_split.h"
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
    // The 8-wide collapse of the same binary tree over the same leaves,
    // already quantized (empty when it would be too deep for the stack).
    std::vector<BvhNode8Q> nodes8;
    uint32_t max_depth8 = 0;
};

// A node of the binary tree both builders grow (40 bytes, the layout the
// device builder writes).
struct BinaryNode {
    Box3 box;
    uint32_t left = 0, right = 0;   // children (internal)
    uint32_t first = 0, count = 0;  // leaf range in the builder's order (count > 0)
};
static_assert(sizeof(BinaryNode) == 40, "device and host share this layout");

// Binned-SAH binary build (32 bins on the three axes, leaves of <= 8
// triangles), collapsed to 4-wide nodes by repeatedly opening the child of
// largest surface area. Returns false with a message when the tree would
// be deeper than the traversal stack allows.
bool build_bvh4(const std::vector<WorldTriangle>& input, uint32_t max_stack, Bvh4* out, std::string* error);

// The collapse step alone, from a binary tree rooted at node 0 (nodes may
// be numbered in any order); max_abs is the scene's largest |coordinate|.
bool collapse_to_bvh4(const std::vector<BinaryNode>& bnodes, float max_abs, uint32_t max_stack, Bvh4* out,
                      std::string* error);

// The 8-wide collapse of the same binary tree, by the same greedy rule up
// to eight children, straight to BvhNode8Q; fills out->nodes8 and
// max_depth8, or leaves nodes8 empty (returning false) when the tree is too
// deep for `max_stack` at seven pushes per level.
bool collapse_to_bvh8(const std::vector<BinaryNode>& bnodes, float max_abs, uint32_t max_stack, Bvh4* out);

// The device-resident outputs of the device builder: device allocations
// the caller takes over (and frees with hipFree); nodes8 is null when the
// 8-wide tree would be too deep for the stack.
struct DeviceBvh {
    BvhTriangle* tris = nullptr;
    BvhNode4Q* nodes = nullptr;
    BvhNode8Q* nodes8 = nullptr;
};

// The same trees built entirely on the device (cn_build.hip): top levels by
// device-wide binning with every split decided on the device, subtrees of
// <= 2048 triangles by one workgroup each, the collapse to 4- and 8-wide
// nodes and their quantization on the device. `out` receives host copies of
// the full-precision 4-wide nodes and the quantized 8-wide nodes (for
// read-back); the traversal's arrays stay on the device in `*dev`.
bool build_bvh_device(const std::vector<WorldTriangle>& input, uint32_t max_stack, Bvh4* out, DeviceBvh* dev,
                      std::string* error);

// Rewrites, on the device, the flags word (v2.w) of every leaf-order
// triangle from its instance's entry in `instance_flags` (host array of
// `instances` words): a material edit changes no bounds, so the tree stays.
bool patch_triangle_flags(BvhTriangle* tris, uint32_t count, const uint32_t* instance_flags, uint32_t instances,
                          std::string* error);

// The packet walk's 128-byte nodes (BvhNode4P) from the device's quantized
// ones, node for node, on the device (the same boxes, planes decoded).
bool decode_packet_nodes(const BvhNode4Q* nodes, uint32_t count, BvhNode4P* out, std::string* error);

// The traversal's 64-byte nodes (BvhNode4Q) from the full-precision ones,
// node for node: each node's grid starts at the low corner of its
// children's union and steps by the smallest power of two that spans the
// union in 255 steps; planes round outward.
void quantize_bvh4(const BvhNode4* nodes, size_t count, BvhNode4Q* out);

}  // namespace cn
