// cn_bvh.cpp — binned-SAH BVH build, collapsed to 4-wide nodes (see cn_bvh.h).

#include "cn_bvh.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

namespace cn {

namespace {

float bits_float(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::fmin(lo[k], b.lo[k]);
            hi[k] = std::fmax(hi[k], b.hi[k]);
        }
    }
    void grow(const float p[3]) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::fmin(lo[k], p[k]);
            hi[k] = std::fmax(hi[k], p[k]);
        }
    }
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Prim {
    Box box;
    float c[3];
};

struct BNode {
    Box box;
    uint32_t left = 0, right = 0;  // children (internal)
    uint32_t first = 0, count = 0;  // leaf range (count > 0)
};

constexpr int BINS = 32;
constexpr uint32_t LEAF_SIZE = 4;

struct Builder {
    const std::vector<Prim>& prims;
    std::vector<uint32_t>& order;
    std::vector<BNode>& nodes;
    std::atomic<uint32_t> next{1};

    Builder(const std::vector<Prim>& p, std::vector<uint32_t>& o, std::vector<BNode>& n)
        : prims(p), order(o), nodes(n) {}

    uint32_t alloc2() { return next.fetch_add(2); }

    void build(uint32_t node, uint32_t begin, uint32_t end, int depth) {
        BNode& out = nodes[node];
        Box box, cbox;
        for (uint32_t i = begin; i < end; ++i) {
            const Prim& p = prims[order[i]];
            box.grow(p.box);
            cbox.grow(p.c);
        }
        out.box = box;
        const uint32_t n = end - begin;
        if (n <= 1) {
            out.first = begin;
            out.count = n;
            return;
        }
        // Binned SAH over all three axes.
        float best_cost = INFINITY;
        int best_axis = -1, best_split = 0;
        for (int axis = 0; axis < 3; ++axis) {
            const float extent = cbox.hi[axis] - cbox.lo[axis];
            if (!(extent > 0.0f)) continue;
            const float scale = float(BINS) / extent;
            Box bins[BINS];
            uint32_t counts[BINS] = {};
            for (uint32_t i = begin; i < end; ++i) {
                const Prim& p = prims[order[i]];
                int b = int((p.c[axis] - cbox.lo[axis]) * scale);
                b = std::min(std::max(b, 0), BINS - 1);
                bins[b].grow(p.box);
                ++counts[b];
            }
            float right_area[BINS];
            uint32_t right_count[BINS];
            Box acc;
            uint32_t cnt = 0;
            for (int b = BINS - 1; b > 0; --b) {
                acc.grow(bins[b]);
                cnt += counts[b];
                right_area[b] = acc.area();
                right_count[b] = cnt;
            }
            Box lacc;
            uint32_t lcnt = 0;
            for (int b = 0; b < BINS - 1; ++b) {
                lacc.grow(bins[b]);
                lcnt += counts[b];
                if (lcnt == 0 || right_count[b + 1] == 0) continue;
                const float cost = lacc.area() * float(lcnt) + right_area[b + 1] * float(right_count[b + 1]);
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_split = b;
                }
            }
        }
        const float parent_area = box.area();
        const float leaf_cost = float(n);
        const float split_cost = 1.0f + (parent_area > 0.0f ? best_cost / parent_area : 0.0f);
        if (n <= LEAF_SIZE && (best_axis < 0 || leaf_cost <= split_cost)) {
            out.first = begin;
            out.count = n;
            return;
        }
        uint32_t mid;
        if (best_axis < 0) {
            // All centroids coincide: split the range in half.
            mid = begin + n / 2;
        } else {
            const float extent = cbox.hi[best_axis] - cbox.lo[best_axis];
            const float scale = float(BINS) / extent;
            const float lo = cbox.lo[best_axis];
            auto* first = order.data() + begin;
            auto* last = order.data() + end;
            auto* m = std::partition(first, last, [&](uint32_t id) {
                int b = int((prims[id].c[best_axis] - lo) * scale);
                b = std::min(std::max(b, 0), BINS - 1);
                return b <= best_split;
            });
            mid = uint32_t(m - order.data());
            if (mid == begin || mid == end) mid = begin + n / 2;
        }
        const uint32_t children = alloc2();
        out.left = children;
        out.right = children + 1;
        out.count = 0;
        if (depth < 6 && n > 4096) {
            std::thread t([&, children, begin, mid, depth]() { build(children, begin, mid, depth + 1); });
            build(children + 1, mid, end, depth + 1);
            t.join();
        } else {
            build(children, begin, mid, depth + 1);
            build(children + 1, mid, end, depth + 1);
        }
    }
};

struct Collapser {
    const std::vector<BNode>& bnodes;
    std::vector<BvhNode4>& out;
    float pad;
    uint32_t max_depth = 0;

    static void set_child(BvhNode4& n, int slot, const Box& b, uint32_t child, float pad) {
        float* lx = &n.lo_x.x;
        float* hx = &n.hi_x.x;
        float* ly = &n.lo_y.x;
        float* hy = &n.hi_y.x;
        float* lz = &n.lo_z.x;
        float* hz = &n.hi_z.x;
        lx[slot] = b.lo[0] - pad;
        hx[slot] = b.hi[0] + pad;
        ly[slot] = b.lo[1] - pad;
        hy[slot] = b.hi[1] + pad;
        lz[slot] = b.lo[2] - pad;
        hz[slot] = b.hi[2] + pad;
        (&n.child.x)[slot] = child;
    }

    static BvhNode4 empty_node() {
        BvhNode4 n;
        n.lo_x = n.lo_y = n.lo_z = make_float4(INFINITY, INFINITY, INFINITY, INFINITY);
        n.hi_x = n.hi_y = n.hi_z = make_float4(-INFINITY, -INFINITY, -INFINITY, -INFINITY);
        n.child = make_uint4(CHILD_EMPTY, CHILD_EMPTY, CHILD_EMPTY, CHILD_EMPTY);
        n.meta = make_uint4(0, 0, 0, 0);
        return n;
    }

    uint32_t encode(uint32_t b, uint32_t depth) {
        const BNode& n = bnodes[b];
        if (n.count > 0) return leaf_encode(n.first, n.count);
        return collapse(b, depth);
    }

    uint32_t collapse(uint32_t b, uint32_t depth) {
        max_depth = std::max(max_depth, depth);
        const uint32_t index = uint32_t(out.size());
        out.push_back(empty_node());
        uint32_t kids[4];
        int count = 0;
        const BNode& root = bnodes[b];
        if (root.count > 0) {
            kids[count++] = b;
        } else {
            kids[count++] = root.left;
            kids[count++] = root.right;
            while (count < 4) {
                int open = -1;
                float best = -1.0f;
                for (int k = 0; k < count; ++k) {
                    const BNode& c = bnodes[kids[k]];
                    if (c.count == 0 && c.box.area() > best) {
                        best = c.box.area();
                        open = k;
                    }
                }
                if (open < 0) break;
                const BNode& o = bnodes[kids[open]];
                kids[open] = o.left;
                kids[count++] = o.right;
            }
        }
        uint32_t encoded[4];
        for (int k = 0; k < count; ++k) encoded[k] = encode(kids[k], depth + 1);
        BvhNode4& node = out[index];
        for (int k = 0; k < count; ++k) set_child(node, k, bnodes[kids[k]].box, encoded[k], pad);
        node.meta.x = depth;
        return index;
    }
};

}  // namespace

bool build_bvh4(const std::vector<WorldTriangle>& input, uint32_t max_stack, Bvh4* out, std::string* error) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = uint32_t(input.size());
    if (n == 0) {
        *error = "a scene needs at least one triangle";
        return false;
    }
    if (n > LEAF_FIRST_MASK) {
        *error = "too many triangles for the BVH leaf encoding";
        return false;
    }
    std::vector<Prim> prims(n);
    float max_abs = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        const WorldTriangle& t = input[i];
        const float p[3][3] = {{t.v0.x, t.v0.y, t.v0.z}, {t.v1.x, t.v1.y, t.v1.z}, {t.v2.x, t.v2.y, t.v2.z}};
        Prim& pr = prims[i];
        for (const auto& q : p) {
            pr.box.grow(q);
            for (int k = 0; k < 3; ++k) max_abs = std::fmax(max_abs, std::fabs(q[k]));
        }
        for (int k = 0; k < 3; ++k) pr.c[k] = 0.5f * (pr.box.lo[k] + pr.box.hi[k]);
    }
    if (!std::isfinite(max_abs)) {
        *error = "scene geometry holds a non-finite vertex";
        return false;
    }
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::vector<BNode> bnodes(2 * size_t(n));
    Builder builder(prims, order, bnodes);
    builder.build(0, 0, n, 0);
    bnodes.resize(builder.next.load());

    // Boxes grow by an absolute margin proportional to the scene's extent so
    // the traversal's fused slab test stays conservative.
    Collapser col{bnodes, out->nodes, 4e-6f * std::fmax(max_abs, 1e-3f)};
    out->nodes.clear();
    out->nodes.reserve(bnodes.size() / 2 + 1);
    col.collapse(0, 1);
    out->max_depth = col.max_depth;
    if (3 * col.max_depth > max_stack) {
        *error = "BVH too deep for the traversal stack (depth " + std::to_string(col.max_depth) + ")";
        return false;
    }

    out->tris.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const WorldTriangle& t = input[order[i]];
        BvhTriangle& r = out->tris[i];
        r.v0 = make_float4(t.v0.x, t.v0.y, t.v0.z, bits_float(t.instance));
        r.v1 = make_float4(t.v1.x, t.v1.y, t.v1.z, bits_float(t.primitive));
        r.v2 = make_float4(t.v2.x, t.v2.y, t.v2.z, bits_float(t.flags));
    }
    out->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

}  // namespace cn
