// cn_bvh.cpp — binned-SAH BVH build, collapsed to 4-wide nodes (see cn_bvh.h).

#include "cn_bvh.h"

#include "cn_bvh_split.h"

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

using Box = Box3;

struct Prim {
    Box box;
    float c[3];
};

using BNode = BinaryNode;

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
        Box box = Box::empty(), cbox = Box::empty();
        for (uint32_t i = begin; i < end; ++i) {
            const Prim& p = prims[order[i]];
            box.grow(p.box);
            for (int k = 0; k < 3; ++k) {
                cbox.lo[k] = std::fmin(cbox.lo[k], p.c[k]);
                cbox.hi[k] = std::fmax(cbox.hi[k], p.c[k]);
            }
        }
        out.box = box;
        const uint32_t n = end - begin;
        Box bins[3][BVH_BINS];
        uint32_t counts[3][BVH_BINS] = {};
        for (int axis = 0; axis < 3; ++axis) {
            for (int b = 0; b < BVH_BINS; ++b) bins[axis][b] = Box::empty();
            if (n <= 1 || !(cbox.hi[axis] - cbox.lo[axis] > 0.0f)) continue;
            const float scale = bvh_scale(cbox.lo[axis], cbox.hi[axis]);
            for (uint32_t i = begin; i < end; ++i) {
                const Prim& p = prims[order[i]];
                const int b = bvh_bin(p.c[axis], cbox.lo[axis], scale);
                bins[axis][b].grow(p.box);
                ++counts[axis][b];
            }
        }
        const SplitDecision d = bvh_decide(box, cbox, bins, counts, n);
        if (d.leaf) {
            out.first = begin;
            out.count = n;
            return;
        }
        // Stable partition, so that the device builder (which scatters with
        // prefix sums) orders every range the same way.
        uint32_t mid = begin + n / 2;
        if (d.axis >= 0) {
            const uint32_t left = bvh_left_count(d, counts);
            if (left > 0 && left < n) {
                const float lo = cbox.lo[d.axis];
                const float scale = bvh_scale(cbox.lo[d.axis], cbox.hi[d.axis]);
                std::stable_partition(order.data() + begin, order.data() + end, [&](uint32_t id) {
                    return bvh_bin(prims[id].c[d.axis], lo, scale) <= d.split;
                });
                mid = begin + left;
            }
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

// The 8-wide collapse: the same greedy opening (the internal child of
// largest surface area, until eight children or only leaves remain), each
// node written straight to its quantized form.
struct Collapser8 {
    const std::vector<BNode>& bnodes;
    std::vector<BvhNode8Q>& out;
    float pad;
    uint32_t max_depth = 0;

    uint32_t encode(uint32_t b, uint32_t depth) {
        const BNode& n = bnodes[b];
        if (n.count > 0) return leaf_encode(n.first, n.count);
        return collapse(b, depth);
    }

    uint32_t collapse(uint32_t b, uint32_t depth) {
        max_depth = std::max(max_depth, depth);
        const uint32_t index = uint32_t(out.size());
        out.emplace_back();
        uint32_t kids[8];
        int count = 0;
        const BNode& root = bnodes[b];
        if (root.count > 0) {
            kids[count++] = b;
        } else {
            kids[count++] = root.left;
            kids[count++] = root.right;
            while (count < 8) {
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
        uint32_t encoded[8];
        Box boxes[8];
        for (int k = 0; k < count; ++k) {
            encoded[k] = encode(kids[k], depth + 1);
            boxes[k] = bnodes[kids[k]].box;
            for (int a = 0; a < 3; ++a) {
                boxes[k].lo[a] -= pad;
                boxes[k].hi[a] += pad;
            }
        }
        out[index] = quantize_wide(boxes, encoded, count, depth);
        return index;
    }

    // One node's grid and byte planes, as quantize_node does for 4 children.
    static BvhNode8Q quantize_wide(const Box* boxes, const uint32_t* child, int count, uint32_t depth) {
        BvhNode8Q q{};
        uint32_t* planes_lo[3] = {q.lo_x, q.lo_y, q.lo_z};
        uint32_t* planes_hi[3] = {q.hi_x, q.hi_y, q.hi_z};
        float origin[3];
        for (int k = 0; k < 3; ++k) {
            double a = INFINITY, b = -INFINITY;
            for (int c = 0; c < count; ++c) {
                a = std::min(a, double(boxes[c].lo[k]));
                b = std::max(b, double(boxes[c].hi[k]));
            }
            if (!(a <= b)) a = b = 0.0;
            origin[k] = float(a);
            const double extent = b - double(origin[k]);
            int e = -126;
            if (extent > 0.0) {
                std::frexp(extent / 255.0, &e);
                e = std::max(e, -126);
                while (std::ldexp(255.0, e) < extent) ++e;
                while (e > -126 && std::ldexp(255.0, e - 1) >= extent) --e;
            }
            const double step = std::ldexp(1.0, e);
            for (int c = 0; c < count; ++c) {
                const uint32_t ql = uint32_t(std::floor((double(boxes[c].lo[k]) - double(origin[k])) / step));
                const uint32_t qh = uint32_t(std::ceil((double(boxes[c].hi[k]) - double(origin[k])) / step));
                planes_lo[k][c / 4] |= std::min(ql, 255u) << (8 * (c % 4));
                planes_hi[k][c / 4] |= std::min(qh, 255u) << (8 * (c % 4));
            }
            q.exps |= uint32_t(e + 127) << (8 * k);
        }
        q.ox = origin[0];
        q.oy = origin[1];
        q.oz = origin[2];
        for (int c = 0; c < 8; ++c) q.child[c] = c < count ? child[c] : CHILD_EMPTY;
        q.depth = depth;
        return q;
    }
};

}  // namespace

bool collapse_to_bvh8(const std::vector<BinaryNode>& bnodes, float max_abs, uint32_t max_stack, Bvh4* out) {
    Collapser8 col{bnodes, out->nodes8, 4e-6f * std::fmax(max_abs, 1e-3f)};
    out->nodes8.clear();
    out->nodes8.reserve(bnodes.size() / 4 + 1);
    col.collapse(0, 1);
    out->max_depth8 = col.max_depth;
    if (7 * col.max_depth > max_stack) {  // too deep for the stack at 7 pushes per level: 4-wide only
        out->nodes8.clear();
        return false;
    }
    return true;
}

bool collapse_to_bvh4(const std::vector<BinaryNode>& bnodes, float max_abs, uint32_t max_stack, Bvh4* out,
                      std::string* error) {
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
    return true;
}

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
    bool finite = true;
    for (uint32_t i = 0; i < n; ++i) {
        const WorldTriangle& t = input[i];
        const float p[3][3] = {{t.v0.x, t.v0.y, t.v0.z}, {t.v1.x, t.v1.y, t.v1.z}, {t.v2.x, t.v2.y, t.v2.z}};
        Prim& pr = prims[i];
        bvh_prim_bounds(p, &pr.box, pr.c);
        for (const auto& q : p)
            for (int k = 0; k < 3; ++k) {
                max_abs = std::fmax(max_abs, std::fabs(q[k]));
                finite = finite && std::isfinite(q[k]);
            }
    }
    if (!finite) {
        *error = "scene geometry holds a non-finite vertex";
        return false;
    }
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::vector<BNode> bnodes(2 * size_t(n));
    Builder builder(prims, order, bnodes);
    builder.build(0, 0, n, 0);
    bnodes.resize(builder.next.load());
    if (!collapse_to_bvh4(bnodes, max_abs, max_stack, out, error)) return false;
    collapse_to_bvh8(bnodes, max_abs, max_stack, out);

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

static BvhNode4Q quantize_node(const BvhNode4& n) {
    const float* lo[3] = {&n.lo_x.x, &n.lo_y.x, &n.lo_z.x};
    const float* hi[3] = {&n.hi_x.x, &n.hi_y.x, &n.hi_z.x};
    const uint32_t* child = &n.child.x;
    BvhNode4Q q{};
    float origin[3], steps[3];
    uint32_t planes_lo[3] = {0, 0, 0}, planes_hi[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        double a = INFINITY, b = -INFINITY;
        for (int c = 0; c < 4; ++c) {
            if (child[c] == CHILD_EMPTY) continue;
            a = std::min(a, double(lo[k][c]));
            b = std::max(b, double(hi[k][c]));
        }
        if (!(a <= b)) a = b = 0.0;  // no children: any grid
        origin[k] = float(a);        // a is a float already: exact
        // The smallest step 2^e (e >= -126, a normal float) with 255 steps
        // covering [origin, b]; origins and planes are floats, so the
        // differences below are exact in double.
        const double extent = b - double(origin[k]);
        int e = -126;
        if (extent > 0.0) {
            std::frexp(extent / 255.0, &e);
            e = std::max(e, -126);
            while (std::ldexp(255.0, e) < extent) ++e;
            while (e > -126 && std::ldexp(255.0, e - 1) >= extent) --e;
        }
        const double step = std::ldexp(1.0, e);
        for (int c = 0; c < 4; ++c) {
            uint32_t ql = 0, qh = 0;
            if (child[c] != CHILD_EMPTY) {
                ql = uint32_t(std::floor((double(lo[k][c]) - double(origin[k])) / step));
                qh = uint32_t(std::ceil((double(hi[k][c]) - double(origin[k])) / step));
            }
            planes_lo[k] |= std::min(ql, 255u) << (8 * c);
            planes_hi[k] |= std::min(qh, 255u) << (8 * c);
        }
        steps[k] = float(step);  // 2^e, -126 <= e <= 127: exact
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
    return q;
}

void quantize_bvh4(const BvhNode4* nodes, size_t count, BvhNode4Q* out) {
    for (size_t i = 0; i < count; ++i) out[i] = quantize_node(nodes[i]);
}

}  // namespace cn
