// cn_traverse.h — BVH4 traversal for the intersect and trace_shadow stages
// (the driver ray-query loops of shaders/intersect.slang:64-88 and
// trace_shadow.slang:50-72, rebuilt for CDNA, which has no ray tracing
// units).
//
// One lane per ray, closest-first descent with a short stack (LDS, spilling
// to HBM), and the speculative while-while schedule of Aila & Laine
// ("Understanding the Efficiency of Ray Traversal on GPUs", HPG 2009): a
// lane that reaches a leaf postpones it and keeps descending until every
// lane of the wave64 holds a leaf, so leaf tests run with full SIMD lanes.
#pragma once

#include <type_traits>

#include "cn_launch.h"
#include "cn_math.h"
#include "cn_records.h"

namespace cn {

constexpr uint32_t NODE_DONE = CHILD_EMPTY;  // never pushed: empty slots fail every box test

// The traversal stack: the first STACK_LDS entries of each lane live in LDS
// (entry k of thread t at [k * TRAVERSAL_BLOCK + t], bank-conflict free),
// the rest in a per-lane spill area in HBM laid out [entry][lane] so spills
// coalesce. The builder rejects trees that could need more than STACK_MAX
// entries; a push past it is dropped all the same, so every pop stays in
// bounds. The LDS part is typed as LDS (address space 3) and the spill as
// global memory (address space 1), so each access is a ds_* or global_*
// instruction; through plain pointers the compiler merged the two loads of
// a pop into one flat load, which waits on every outstanding global load
// too. The top is kept as the LDS byte offset of the lane's next free entry
// (entry count * STACK_ROW + threadIdx.x * 4), so an LDS access needs no
// address arithmetic of its own.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) char lds_char;
typedef __attribute__((address_space(1))) uint32_t global_u32;

constexpr uint32_t STACK_ROW = TRAVERSAL_BLOCK * 4;  // bytes between a lane's consecutive LDS entries
constexpr uint32_t STACK_LDS_END = STACK_LDS * STACK_ROW;

struct Stack {
    lds_u32* lds;
    global_u32* spill;
    uint32_t lanes;
    uint32_t lane;
    uint32_t base;  // this lane's entry 0: threadIdx.x * 4 (the stack is empty when at == base)
    uint32_t at;    // byte offset of the next free entry
    __device__ __forceinline__ lds_u32& slot(uint32_t offset) {
        return *reinterpret_cast<lds_u32*>(reinterpret_cast<lds_char*>(lds) + offset);
    }
    __device__ __forceinline__ global_u32& spilled(uint32_t offset) {
        return spill[((offset - base) / STACK_ROW - STACK_LDS) * lanes + lane];
    }
    __device__ __forceinline__ void clear() { at = base; }
    __device__ __forceinline__ void push(uint32_t v) {
        if (at - base >= STACK_MAX * STACK_ROW) return;
        if (at < STACK_LDS_END) slot(at) = v;
        else spilled(at) = v;
        at += STACK_ROW;
    }
    // Pushes the hit children after the nearest, farthest first: c3, c2,
    // c1 for nh = 4, c2, c1 for 3, c1 for 2 (nh >= 1). Away from the LDS
    // boundary the three stores go out unconditionally, c3 at the top, c2
    // and c1 at the top plus (nh - 3) and (nh - 2) entries clamped at 0:
    // a child past the hit count lands on the top before a hit child
    // overwrites it, or above the new top, where nothing reads.
    __device__ __forceinline__ void push_sorted(uint32_t c3, uint32_t c2, uint32_t c1, uint32_t nh) {
        if (at < (STACK_LDS - 2) * STACK_ROW) {  // entries top .. top + 2 are in LDS
            const uint32_t n = nh * STACK_ROW;
            const uint32_t up2 = n > 3 * STACK_ROW ? n - 3 * STACK_ROW : 0u;
            const uint32_t up1 = n > 2 * STACK_ROW ? n - 2 * STACK_ROW : 0u;
            slot(at) = c3;
            slot(at + up2) = c2;
            slot(at + up1) = c1;
            at += n - STACK_ROW;
            return;
        }
        if (nh > 3) push(c3);
        if (nh > 2) push(c2);
        if (nh > 1) push(c1);
    }
    // The 8-wide form: ref[1..7] pushed farthest first, the ones past the
    // hit count landing above the new top.
    __device__ __forceinline__ void push_sorted8(const uint32_t (&ref)[8], uint32_t nh) {
        if (at < (STACK_LDS - 6) * STACK_ROW) {  // entries top .. top + 6 are in LDS
            uint32_t a = at;
#pragma unroll
            for (int k = 7; k >= 1; --k) {
                slot(a) = ref[k];
                a += uint32_t(k) < nh ? STACK_ROW : 0u;
            }
            at = a;
            return;
        }
#pragma unroll
        for (int k = 7; k >= 1; --k)
            if (uint32_t(k) < nh) push(ref[k]);
    }
    __device__ __forceinline__ uint32_t pop_or_done() {
        if (at == base) return NODE_DONE;
        at -= STACK_ROW;
        return at < STACK_LDS_END ? uint32_t(slot(at)) : uint32_t(spilled(at));
    }
};

__device__ __forceinline__ Stack make_stack(uint32_t* lds_stack, uint32_t* spill, uint32_t lanes, uint32_t lane) {
    const uint32_t base = threadIdx.x * 4;
    return Stack{(lds_u32*)lds_stack, (global_u32*)spill, lanes, lane, base, base};
}

__device__ __forceinline__ float safe_rcp(float x) {
    return 1.0f / (fabsf(x) < 1e-20f ? copysignf(1e-20f, x) : x);
}

__device__ __forceinline__ float pick(v3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Watertight ray/triangle test (Woop, Benthin & Wald, JCGT 2(1), 2013),
// restated identically by the CPU oracle so both agree on the hit bit for
// bit: t in [0, tmax], barycentrics (b1, b2) on v1 and v2.
struct WRay {
    v3 o;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

__device__ __forceinline__ WRay make_wray(v3 o, v3 d) {
    WRay w;
    w.o = o;
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = 0;
    float m = ax;
    if (ay > m) {
        kz = 1;
        m = ay;
    }
    if (az > m) kz = 2;
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    const float dz = pick(d, kz);
    if (dz < 0.0f) {
        const int t = kx;
        kx = ky;
        ky = t;
    }
    w.kx = kx;
    w.ky = ky;
    w.kz = kz;
    w.Sx = pick(d, kx) / dz;
    w.Sy = pick(d, ky) / dz;
    w.Sz = 1.0f / dz;
    return w;
}

__device__ __forceinline__ bool hit_triangle(const WRay& w, v3 v0, v3 v1, v3 v2, float tmax, float* t_out,
                                             float* b1, float* b2) {
    const v3 A = v0 - w.o, B = v1 - w.o, C = v2 - w.o;
    float Az = pick(A, w.kz), Bz = pick(B, w.kz), Cz = pick(C, w.kz);
    const float Ax = pick(A, w.kx) - w.Sx * Az, Ay = pick(A, w.ky) - w.Sy * Az;
    const float Bx = pick(B, w.kx) - w.Sx * Bz, By = pick(B, w.ky) - w.Sy * Bz;
    const float Cx = pick(C, w.kx) - w.Sx * Cz, Cy = pick(C, w.ky) - w.Sy * Cz;
    float U = Cx * By - Cy * Bx;
    float Vv = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || Vv == 0.0f || W == 0.0f) {
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        Vv = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.0f || Vv < 0.0f || W < 0.0f) && (U > 0.0f || Vv > 0.0f || W > 0.0f)) return false;
    const float det = U + Vv + W;
    if (det == 0.0f) return false;
    Az = w.Sz * Az;
    Bz = w.Sz * Bz;
    Cz = w.Sz * Cz;
    const float T = U * Az + Vv * Bz + W * Cz;
    const float rcp = 1.0f / det;
    const float t = T * rcp;
    if (!(t >= 0.0f) || !(t <= tmax)) return false;
    *t_out = t;
    *b1 = Vv * rcp;
    *b2 = W * rcp;
    return true;
}

struct SlabRay {
    float ix, iy, iz, x, y, z;  // reciprocal direction, origin
    bool nx, ny, nz;            // the axis's reciprocal is negative: the near plane is the high one
};

__device__ __forceinline__ SlabRay make_slab(v3 o, v3 d) {
    SlabRay r;
    r.ix = safe_rcp(d.x);
    r.iy = safe_rcp(d.y);
    r.iz = safe_rcp(d.z);
    r.x = o.x;
    r.y = o.y;
    r.z = o.z;
    r.nx = r.ix < 0.0f;
    r.ny = r.iy < 0.0f;
    r.nz = r.iz < 0.0f;
    return r;
}

// Byte c of a plane word as a float (one v_cvt_f32_ubyte).
__device__ __forceinline__ float plane(uint32_t word, int c) { return float((word >> (8 * c)) & 0xffu); }

// One compare-exchange by key, swapping on strictly greater.
__device__ __forceinline__ void order_pair(float& ka, uint32_t& ca, float& kb, uint32_t& cb) {
    const bool swap = ka > kb;
    const float k = swap ? kb : ka;
    const uint32_t c = swap ? cb : ca;
    kb = swap ? ka : kb;
    cb = swap ? ca : cb;
    ka = k;
    ca = c;
}

// Box-test the four children of one node against [0, bound]; push the hit
// ones farthest first and return the nearest (or the next stack entry).
// Hit children are ordered by entry distance with the optimal four-input
// sorting network (five compare-exchanges in three rounds, no branches),
// misses keyed +inf so they sort last. Equal keys may leave slot order,
// which changes only the visiting order: the closest hit, visibility and
// transmittance do not depend on it.
// A plane at origin + q * step meets the ray at t = q * (step / d) +
// (origin - o) / d: one fma per plane after six per-node products. The
// step is positive, so on each axis the low plane is met first exactly
// when d is positive: the near and far plane words are picked once per node
// by the ray's signs, and the per-child test needs no min/max per axis
// (the same values the min/max form gives; a child's low byte never
// exceeds its high one). The boxes' padding (far above these roundings)
// keeps the test conservative.
__device__ __forceinline__ uint32_t visit(const BvhNode4Q* __restrict__ nodes, uint32_t index, const SlabRay& r,
                                          float bound, Stack& st) {
    // A 32-bit byte offset from the tree's base (a node index is below 2^26:
    // leaves address at most 2^27 triangles, a node holds four children),
    // so the fetch takes the base from scalar registers.
    const BvhNode4Q n =
        *reinterpret_cast<const BvhNode4Q*>(reinterpret_cast<const char*>(nodes) + index * uint32_t(sizeof(BvhNode4Q)));
    const float sx = n.step_x * r.ix;
    const float sy = n.step_y * r.iy;
    const float sz = n.step_z * r.iz;
    const float bx0 = (n.ox - r.x) * r.ix, by0 = (n.oy - r.y) * r.iy, bz0 = (n.oz - r.z) * r.iz;
    const uint32_t near_x = r.nx ? n.hi_x : n.lo_x, far_x = r.nx ? n.lo_x : n.hi_x;
    const uint32_t near_y = r.ny ? n.hi_y : n.lo_y, far_y = r.ny ? n.lo_y : n.hi_y;
    const uint32_t near_z = r.nz ? n.hi_z : n.lo_z, far_z = r.nz ? n.lo_z : n.hi_z;
    const uint32_t ch[4] = {n.child.x, n.child.y, n.child.z, n.child.w};
    float key[4];
    uint32_t ref[4];
    uint32_t nh = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float ax = fmaf(plane(near_x, c), sx, bx0), bx = fmaf(plane(far_x, c), sx, bx0);
        const float ay = fmaf(plane(near_y, c), sy, by0), by = fmaf(plane(far_y, c), sy, by0);
        const float az = fmaf(plane(near_z, c), sz, bz0), bz = fmaf(plane(far_z, c), sz, bz0);
        const float t0 = fmaxf(fmaxf(ax, ay), fmaxf(az, 0.0f));
        const float t1 = fminf(fminf(bx, by), fminf(bz, bound)) * 1.0000004f;
        const bool hit = t0 <= t1 && ch[c] != CHILD_EMPTY;
        key[c] = hit ? t0 : __int_as_float(0x7f800000);
        ref[c] = hit ? ch[c] : NODE_DONE;
        nh += hit ? 1u : 0u;
    }
    order_pair(key[0], ref[0], key[1], ref[1]);
    order_pair(key[2], ref[2], key[3], ref[3]);
    order_pair(key[0], ref[0], key[2], ref[2]);
    order_pair(key[1], ref[1], key[3], ref[3]);
    order_pair(key[1], ref[1], key[2], ref[2]);
    if (nh == 0) return st.pop_or_done();
    st.push_sorted(ref[3], ref[2], ref[1], nh);
    return ref[0];
}

// The same box test over an 8-wide node (six dwordx4 fetches), its hit
// children ordered by entry distance with Batcher's odd-even merge network
// (19 compare-exchanges in six rounds; misses keyed +inf sort last). Ties
// may leave slot order, which changes only the visiting order: the closest
// hit, visibility and transmittance do not depend on it.
__device__ __forceinline__ uint32_t visit(const BvhNode8Q* __restrict__ nodes, uint32_t index, const SlabRay& r,
                                          float bound, Stack& st) {
    // a 32-bit byte offset, as for the 4-wide nodes (8-wide indices are below 2^25)
    const uint4* line =
        reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(nodes) + index * uint32_t(sizeof(BvhNode8Q)));
    const uint4 h = line[0];
    const uint4 px = line[1];  // lo_x[0], lo_x[1], hi_x[0], hi_x[1]
    const uint4 py = line[2];
    const uint4 pz = line[3];
    const uint4 ca = line[4];
    const uint4 cb = line[5];
    const float sx = __uint_as_float((h.w & 0xffu) << 23) * r.ix;
    const float sy = __uint_as_float(((h.w >> 8) & 0xffu) << 23) * r.iy;
    const float sz = __uint_as_float(((h.w >> 16) & 0xffu) << 23) * r.iz;
    const float bx0 = (__uint_as_float(h.x) - r.x) * r.ix;
    const float by0 = (__uint_as_float(h.y) - r.y) * r.iy;
    const float bz0 = (__uint_as_float(h.z) - r.z) * r.iz;
    const uint32_t nx[2] = {r.nx ? px.z : px.x, r.nx ? px.w : px.y}, fx[2] = {r.nx ? px.x : px.z, r.nx ? px.y : px.w};
    const uint32_t ny[2] = {r.ny ? py.z : py.x, r.ny ? py.w : py.y}, fy[2] = {r.ny ? py.x : py.z, r.ny ? py.y : py.w};
    const uint32_t nz[2] = {r.nz ? pz.z : pz.x, r.nz ? pz.w : pz.y}, fz[2] = {r.nz ? pz.x : pz.z, r.nz ? pz.y : pz.w};
    const uint32_t ch[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
    float key[8];
    uint32_t ref[8];
    uint32_t nh = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int w = c >> 2, b = c & 3;
        const float ax = fmaf(plane(nx[w], b), sx, bx0), bx = fmaf(plane(fx[w], b), sx, bx0);
        const float ay = fmaf(plane(ny[w], b), sy, by0), by = fmaf(plane(fy[w], b), sy, by0);
        const float az = fmaf(plane(nz[w], b), sz, bz0), bz = fmaf(plane(fz[w], b), sz, bz0);
        const float t0 = fmaxf(fmaxf(ax, ay), fmaxf(az, 0.0f));
        const float t1 = fminf(fminf(bx, by), fminf(bz, bound)) * 1.0000004f;
        const bool hit = t0 <= t1 && ch[c] != CHILD_EMPTY;
        key[c] = hit ? t0 : __int_as_float(0x7f800000);
        ref[c] = hit ? ch[c] : NODE_DONE;
        nh += hit ? 1u : 0u;
    }
    if (nh == 0) return st.pop_or_done();
#define CN_CE(a, b) order_pair(key[a], ref[a], key[b], ref[b])
    CN_CE(0, 1); CN_CE(2, 3); CN_CE(4, 5); CN_CE(6, 7);
    CN_CE(0, 2); CN_CE(1, 3); CN_CE(4, 6); CN_CE(5, 7);
    CN_CE(1, 2); CN_CE(5, 6);
    CN_CE(0, 4); CN_CE(1, 5); CN_CE(2, 6); CN_CE(3, 7);
    CN_CE(2, 4); CN_CE(3, 5);
    CN_CE(1, 2); CN_CE(3, 4); CN_CE(5, 6);
#undef CN_CE
    st.push_sorted8(ref, nh);
    return ref[0];
}

// The tree a traversal walks: the 4-wide nodes, or the 8-wide ones.
template <bool W8>
__device__ __forceinline__ auto tree_of(const SceneView& s) {
    if constexpr (W8) return s.nodes8;
    else return s.nodes;
}

// The speculative while-while walk over [0, bound]. `leaf(ref)` tests one
// leaf's triangles (it may shrink `bound`) and returns true to end the ray.
template <typename Node, typename Leaf>
__device__ __forceinline__ void walk(const Node* __restrict__ nodes, const SlabRay& r, float& bound, Stack& st,
                                     uint32_t& nodes_visited, Leaf&& leaf_fn) {
    st.clear();
    uint32_t cur = 0;  // the root is always an internal node
    uint32_t leaf = NODE_DONE;
    while (true) {
        while (!(cur & CHILD_LEAF)) {
            cur = visit(nodes, cur, r, bound, st);
            ++nodes_visited;
            if ((cur & CHILD_LEAF) && cur != NODE_DONE && leaf == NODE_DONE) {
                leaf = cur;
                cur = st.pop_or_done();
            }
            if (!__any(leaf == NODE_DONE)) break;  // every lane still here holds a leaf
        }
        while (leaf != NODE_DONE) {
            if (leaf_fn(leaf)) return;
            leaf = NODE_DONE;
            if ((cur & CHILD_LEAF) && cur != NODE_DONE) {
                leaf = cur;
                cur = st.pop_or_done();
            }
        }
        if (cur == NODE_DONE) return;
    }
}

__device__ __forceinline__ void leaf_range(uint32_t ref, uint32_t* first, uint32_t* count) {
    *first = ref & LEAF_FIRST_MASK;
    *count = ((ref >> 27) & 15u) + 1u;
}

// The packet walk reads the tree and triangles at wave-uniform addresses
// through the scalar unit, once per wave. A packet node (128 bytes) comes in
// one round trip: two 64-byte loads from one base, waited for in the same
// block, so no load is in flight when the registers are handed back.
typedef uint32_t sgpr_x16 __attribute__((ext_vector_type(16)));
__device__ __forceinline__ BvhNode4P load_packet_node(const BvhNode4P* nodes, uint32_t index) {
    const uint64_t at = reinterpret_cast<uint64_t>(nodes + index);
    sgpr_x16 a, b;
    asm volatile(
        "s_load_dwordx16 %0, %2, 0x0\n\t"
        "s_load_dwordx16 %1, %2, 0x40\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&s"(a), "=&s"(b)
        : "s"(at));
    BvhNode4P n;
    uint32_t* w = reinterpret_cast<uint32_t*>(&n);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        w[k] = a[k];
        w[16 + k] = b[k];
    }
    return n;
}
// The scene's triangles as the leaf test reads them: per thread, or through
// the scalar unit at a wave-uniform index.
struct TrisPerLane {
    const BvhTriangle* tris;
    __device__ __forceinline__ BvhTriangle operator[](uint32_t i) const { return tris[i]; }
};
// A triangle (48 bytes at a wave-uniform index) in one scalar round trip,
// waited for in the same block as the node fetch is.
typedef uint32_t sgpr_x8 __attribute__((ext_vector_type(8)));
typedef uint32_t sgpr_x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ BvhTriangle load_packet_triangle(const BvhTriangle* tris, uint32_t index) {
    const uint64_t at = reinterpret_cast<uint64_t>(tris + index);
    sgpr_x8 a;
    sgpr_x4 b;
    asm volatile(
        "s_load_dwordx8 %0, %2, 0x0\n\t"
        "s_load_dwordx4 %1, %2, 0x20\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&s"(a), "=&s"(b)
        : "s"(at));
    BvhTriangle t;
    uint32_t* w = reinterpret_cast<uint32_t*>(&t);
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = a[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[8 + k] = b[k];
    return t;
}
struct TrisUniform {
    const BvhTriangle* tris;
    __device__ __forceinline__ BvhTriangle operator[](uint32_t i) const { return load_packet_triangle(tris, i); }
};

struct TraceHit {
    bool hit;
    float t;
    uint32_t inst, prim;
    float b1, b2;
    uint32_t flags;  // the hit triangle's flags (mask, opacity, material class)
};

// Closest hit among triangles whose instance mask meets `mask`. A crossing
// of a fractional-opacity instance commits when its own Bernoulli trial
// passes (crossing_uniform(pixel, sample, bounce, instance, primitive) <
// opacity; intersect.slang:64-76), so the hit is the nearest crossing whose
// trial commits, in any traversal order. The reference also commits every
// crossing past its eighth in traversal order (intersect.slang:20-23); with
// per-surface trials the walk needs no such bound, and without it the
// result does not depend on the order. Equal-t ties go to the lower
// (instance, primitive).
// (`tris` reads the scene's triangles: TrisPerLane or TrisUniform.)
template <typename Tris>
__device__ __forceinline__ void closest_leaf(const SceneView& s, Tris tris, const WRay& w, uint32_t ref,
                                             uint32_t mask, uint32_t pixel, uint32_t sample, uint32_t bounce,
                                             TraceHit& best, uint32_t& tris_tested) {
    uint32_t first, count;
    leaf_range(ref, &first, &count);
    for (uint32_t k = 0; k < count; ++k) {
        const BvhTriangle tri = tris[first + k];
        ++tris_tested;
        const uint32_t flags = __float_as_uint(tri.v2.w);
        if (((flags >> 8) & mask) == 0) continue;
        float t, b1, b2;
        if (!hit_triangle(w, xyz(tri.v0), xyz(tri.v1), xyz(tri.v2), best.t, &t, &b1, &b2)) continue;
        const uint32_t inst = __float_as_uint(tri.v0.w), prim = __float_as_uint(tri.v1.w);
        if (best.hit && t == best.t && (inst > best.inst || (inst == best.inst && prim > best.prim))) continue;
        if ((flags & TRI_NON_OPAQUE) &&
            !(crossing_uniform(pixel, sample, bounce, inst, prim) < s.materials[inst].opacity))
            continue;
        best = TraceHit{true, t, inst, prim, b1, b2, flags};
    }
}

template <bool W8>
__device__ __forceinline__ TraceHit trace_closest(const SceneView& s, v3 o, v3 d, float tmax, uint32_t mask,
                                                  Stack& st, uint32_t pixel, uint32_t sample, uint32_t bounce,
                                                  uint32_t& nodes_visited, uint32_t& tris_tested) {
    const SlabRay r = make_slab(o, d);
    const WRay w = make_wray(o, d);
    TraceHit best{false, tmax, 0u, 0u, 0.0f, 0.0f, 0u};
    walk(tree_of<W8>(s), r, best.t, st, nodes_visited, [&](uint32_t ref) {
        closest_leaf(s, TrisPerLane{s.tris}, w, ref, mask, pixel, sample, bounce, best, tris_tested);
        return false;
    });
    return best;
}

// Camera rays of one wave walk the 4-wide tree together. The node to visit
// and the stack are the wave's own (one value for all lanes, the stack in
// the wave's LDS rows); each lane tests its ray against the node's four
// children, and the wave descends into every child that some lane's ray
// enters within its closest hit so far, nearest first by the entry
// distances of the first lane that enters any. Each lane thus visits every
// node its own walk would visit (and perhaps more), so its closest hit —
// which does not depend on the visiting order — is the same. Rays through
// one or two pixels enter nearly the same nodes, so the wave walks about
// one ray's path and fetches each node and triangle once for all lanes,
// with no per-lane stack or child sorting. The wave reads the decoded
// nodes (BvhNode4P): the quantized node's boxes with every plane as an f32
// offset, met with one fma per plane — the per-lane walk's test rounded
// once less, conservative within the same outward padding. `wstack` holds
// PACKET_STACK_ROW entries for this wave.
__device__ __forceinline__ uint32_t wave_uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// SIGNS: bit k set when every lane's direction is negative on axis k (its
// near plane on that axis is the high one), or PACKET_MIXED_SIGNS when the
// lanes differ on some axis and each orders every plane pair itself.
constexpr int PACKET_MIXED_SIGNS = 8;
// Entries of a wave's packet stack row: STACK_MAX and the spare entries a
// block push may write above the top.
constexpr uint32_t PACKET_STACK_ROW = STACK_MAX + 4;
template <int SIGNS>
__device__ __forceinline__ TraceHit trace_packet_signs(const SceneView& s, const SlabRay& r, const WRay& w,
                                                       uint32_t mask, uint32_t pixel, uint32_t sample,
                                                       lds_u32* wstack, uint32_t& nodes_visited,
                                                       uint32_t& tris_tested) {
    constexpr bool MIXED = SIGNS == PACKET_MIXED_SIGNS;
    constexpr bool NX = !MIXED && (SIGNS & 1), NY = !MIXED && (SIGNS & 2), NZ = !MIXED && (SIGNS & 4);
    TraceHit best{false, 1e30f, 0u, 0u, 0.0f, 0.0f, 0u};
    const uint32_t leader = __ffsll((unsigned long long)__ballot(1)) - 1;  // the stack's writer
    uint32_t sp = 0;   // wave-uniform
    uint32_t cur = 0;  // wave-uniform; the root is always an internal node
    while (cur != NODE_DONE) {
        if (cur & CHILD_LEAF) {
            closest_leaf(s, TrisUniform{s.tris}, w, cur, mask, pixel, sample, 0, best, tris_tested);
            cur = sp == 0 ? NODE_DONE : wave_uniform(wstack[--sp]);
            continue;
        }
        ++nodes_visited;
        const BvhNode4P n = load_packet_node(s.pnodes, cur);
        const float bx0 = (n.ox - r.x) * r.ix, by0 = (n.oy - r.y) * r.iy, bz0 = (n.oz - r.z) * r.iz;
        const float* near_x = NX ? n.hi_x : n.lo_x;
        const float* far_x = NX ? n.lo_x : n.hi_x;
        const float* near_y = NY ? n.hi_y : n.lo_y;
        const float* far_y = NY ? n.lo_y : n.hi_y;
        const float* near_z = NZ ? n.hi_z : n.lo_z;
        const float* far_z = NZ ? n.lo_z : n.hi_z;
        uint64_t entered[4];
        uint32_t key[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float ax = fmaf(near_x[c], r.ix, bx0), bx = fmaf(far_x[c], r.ix, bx0);
            float ay = fmaf(near_y[c], r.iy, by0), by = fmaf(far_y[c], r.iy, by0);
            float az = fmaf(near_z[c], r.iz, bz0), bz = fmaf(far_z[c], r.iz, bz0);
            if (MIXED) {
                const float lx = ax, ly = ay, lz = az;
                ax = fminf(lx, bx);
                bx = fmaxf(lx, bx);
                ay = fminf(ly, by);
                by = fmaxf(ly, by);
                az = fminf(lz, bz);
                bz = fmaxf(lz, bz);
            }
            const float t0 = fmaxf(fmaxf(ax, ay), fmaxf(az, 0.0f));
            const float t1 = fminf(fminf(bx, by), fminf(bz, best.t)) * 1.0000004f;
            // the lanes whose ray enters the child (ordered t0 <= t1: a NaN
            // never enters), taken straight from the compare; an empty slot's
            // inverted box is never entered when the near planes are picked
            // by sign, and is masked out when each lane orders its planes
            const bool hit = t0 <= t1;
            const uint64_t in = __builtin_amdgcn_fcmpf(t0, t1, 5 /* ordered <= */);
            entered[c] = !MIXED || n.child[c] != CHILD_EMPTY ? in : 0;
            key[c] = hit ? __float_as_uint(t0) : 0x7f800000u;  // t0 >= 0: its bits order as the values
        }
        const uint64_t any = entered[0] | entered[1] | entered[2] | entered[3];
        if (any == 0) {
            cur = sp == 0 ? NODE_DONE : wave_uniform(wstack[--sp]);
            continue;
        }
        // One child entered (the common case for a coherent packet): visit
        // it, with nothing to order or push.
        const uint32_t entered_count = (entered[0] != 0) + (entered[1] != 0) + (entered[2] != 0) + (entered[3] != 0);
        if (entered_count == 1) {
            cur = entered[0] ? n.child[0] : entered[1] ? n.child[1] : entered[2] ? n.child[2] : n.child[3];
            continue;
        }
        // The children some lane enters, keyed by the first such lane's
        // entry distances (those it does not enter after, absent ones last).
        const uint32_t lead = __ffsll((unsigned long long)any) - 1;
        uint32_t k[4], ref[4];
        uint32_t nh = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bool used = entered[c] != 0;
            const uint32_t lead_key = __builtin_amdgcn_readlane(key[c], lead);
            k[c] = used ? lead_key : 0xffffffffu;
            ref[c] = used ? n.child[c] : NODE_DONE;
            nh += used ? 1u : 0u;
        }
        if (nh == 2) {
            // Two children entered: one compare orders them (an unused slot
            // has key ~0, an entered one never does), the farther one is
            // pushed alone.
            const uint32_t ka = k[0] != ~0u ? k[0] : k[1] != ~0u ? k[1] : k[2];
            const uint32_t ra = k[0] != ~0u ? ref[0] : k[1] != ~0u ? ref[1] : ref[2];
            const uint32_t kb = k[3] != ~0u ? k[3] : k[2] != ~0u ? k[2] : k[1];
            const uint32_t rb = k[3] != ~0u ? ref[3] : k[2] != ~0u ? ref[2] : ref[1];
            const bool a_first = ka <= kb;
            if (sp < STACK_MAX) {
                if (__lane_id() == leader) wstack[sp] = a_first ? rb : ra;
                ++sp;
            }
            cur = wave_uniform(a_first ? ra : rb);
            continue;
        }
#define CN_UCE(a, b)                                          \
    do {                                                      \
        const bool swap = k[a] > k[b];                        \
        const uint32_t kk = swap ? k[b] : k[a], rr = swap ? ref[b] : ref[a]; \
        k[b] = swap ? k[a] : k[b];                            \
        ref[b] = swap ? ref[a] : ref[b];                      \
        k[a] = kk;                                            \
        ref[a] = rr;                                          \
    } while (0)
        CN_UCE(0, 1);
        CN_UCE(2, 3);
        CN_UCE(0, 2);
        CN_UCE(1, 3);
        CN_UCE(1, 2);
#undef CN_UCE
        // Push the rest farthest first, as one block: entries sp .. sp + nh - 2
        // get ref[nh - 1] .. ref[1] (the block's other entries lie above the
        // new top, in the row's spare entries at worst). Pushes past
        // STACK_MAX are dropped, as the per-lane stack drops them; the
        // builder's bound keeps it from happening.
        if (nh > 1) {
            const uint32_t p0 = nh == 4 ? ref[3] : nh == 3 ? ref[2] : ref[1];
            const uint32_t p1 = nh == 4 ? ref[2] : ref[1];
            if (__lane_id() == leader) {
                wstack[sp] = p0;
                wstack[sp + 1] = p1;
                wstack[sp + 2] = ref[1];
            }
            sp = min(sp + nh - 1, STACK_MAX);
        }
        cur = wave_uniform(ref[0]);
    }
    return best;
}

__device__ __forceinline__ TraceHit trace_packet(const SceneView& s, v3 o, v3 d, uint32_t mask, uint32_t pixel,
                                                 uint32_t sample, lds_u32* wstack, uint32_t& nodes_visited,
                                                 uint32_t& tris_tested) {
    const SlabRay r = make_slab(o, d);
    const WRay w = make_wray(o, d);
    // Whether every lane's ray points the same way on each axis (nearly
    // always, for the samples of one or two pixels): then one plane of each
    // pair is every lane's near plane, fixed for the whole walk.
    const uint64_t live = __ballot(1);
    const uint64_t sx = __ballot(r.nx), sy = __ballot(r.ny), sz = __ballot(r.nz);
    const bool same = (sx == 0 || sx == live) && (sy == 0 || sy == live) && (sz == 0 || sz == live);
    const int signs = same ? int(sx != 0) | int(sy != 0) << 1 | int(sz != 0) << 2 : PACKET_MIXED_SIGNS;
    switch (signs) {
        case 0: return trace_packet_signs<0>(s, r, w, mask, pixel, sample, wstack, nodes_visited, tris_tested);
        case 1: return trace_packet_signs<1>(s, r, w, mask, pixel, sample, wstack, nodes_visited, tris_tested);
        case 2: return trace_packet_signs<2>(s, r, w, mask, pixel, sample, wstack, nodes_visited, tris_tested);
        case 3: return trace_packet_signs<3>(s, r, w, mask, pixel, sample, wstack, nodes_visited, tris_tested);
        case 4: return trace_packet_signs<4>(s, r, w, mask, pixel, sample, wstack, nodes_visited, tris_tested);
        case 5: return trace_packet_signs<5>(s, r, w, mask, pixel, sample, wstack, nodes_visited, tris_tested);
        case 6: return trace_packet_signs<6>(s, r, w, mask, pixel, sample, wstack, nodes_visited, tris_tested);
        case 7: return trace_packet_signs<7>(s, r, w, mask, pixel, sample, wstack, nodes_visited, tris_tested);
        default:
            return trace_packet_signs<PACKET_MIXED_SIGNS>(s, r, w, mask, pixel, sample, wstack, nodes_visited,
                                                          tris_tested);
    }
}

// The shadow leaf test: returns true when a solid non-target hit nearer
// than t_early ends the ray.
__device__ __forceinline__ bool visibility_leaf(const SceneView& s, const WRay& w, uint32_t ref, uint32_t target_inst,
                                                uint32_t target_prim, float t_early, TraceHit& best,
                                                bool& translucent, uint32_t& tris_tested) {
    uint32_t first, count;
    leaf_range(ref, &first, &count);
    for (uint32_t k = 0; k < count; ++k) {
        const BvhTriangle tri = s.tris[first + k];
        ++tris_tested;
        float t, b1, b2;
        if (!hit_triangle(w, xyz(tri.v0), xyz(tri.v1), xyz(tri.v2), best.t, &t, &b1, &b2)) continue;
        const uint32_t inst = __float_as_uint(tri.v0.w), prim = __float_as_uint(tri.v1.w);
        if (best.hit && t == best.t && (inst > best.inst || (inst == best.inst && prim > best.prim))) continue;
        if (!(inst == target_inst && prim == target_prim)) {
            if (__float_as_uint(tri.v2.w) & TRI_NON_OPAQUE) {
                translucent = true;
                continue;
            }
            if (t < t_early) return true;
        }
        best = TraceHit{true, t, inst, prim, b1, b2, 0u};
    }
    return false;
}

template <bool W8>
__device__ inline float shadow_attenuation(const SceneView& s, const SlabRay& r, const WRay& w, float t_end,
                                           uint32_t target_inst, uint32_t target_prim, Stack& st,
                                           uint32_t& nodes_visited);

// A next-event connection's visibility (trace_shadow.slang:41-71): visible
// iff nothing solid commits within tmax or the committed hit is the target
// (instance, primitive). Returns the attenuation (0 when occluded or
// attenuated below 1e-4): the transmittance of every fractional-opacity
// crossing in front of the connection's end, taken in an order-free form
// (exp2_fixed of a sum of fixed-point logarithms). A solid non-target hit
// nearer than `t_early` occludes without looking further.
template <bool W8>
__device__ __forceinline__ float trace_visibility(const SceneView& s, v3 o, v3 d, float tmax, uint32_t target_inst,
                                                  uint32_t target_prim, float t_early, Stack& st,
                                                  uint32_t& nodes_visited, uint32_t& tris_tested) {
    const SlabRay r = make_slab(o, d);
    const WRay w = make_wray(o, d);
    TraceHit best{false, tmax, 0u, 0u, 0.0f, 0.0f, 0u};
    bool occluded = false, translucent = false;
    walk(tree_of<W8>(s), r, best.t, st, nodes_visited, [&](uint32_t ref) {
        occluded = visibility_leaf(s, w, ref, target_inst, target_prim, t_early, best, translucent, tris_tested);
        return occluded;
    });
    if (occluded) return 0.0f;
    if (best.hit && !(best.inst == target_inst && best.prim == target_prim)) return 0.0f;
    if (!translucent) return 1.0f;
    return shadow_attenuation<W8>(s, r, w, best.hit ? best.t : tmax, target_inst, target_prim, st, nodes_visited);
}

// The rare path of a visible connection that crossed fractional-opacity
// geometry: one more walk over [0, t_end) adds each non-target crossing's
// fixed-point log2 transmittance (the instance record's `log2_transmittance`,
// rounded on the host) and stops once the sum is below the cutoff. Integer
// adds make the result independent of the walk's order.
template <bool W8>
__device__ inline float shadow_attenuation(const SceneView& s, const SlabRay& r, const WRay& w, float t_end,
                                           uint32_t target_inst, uint32_t target_prim, Stack& st,
                                           uint32_t& nodes_visited) {
    int32_t log_sum = 0;
    float bound = t_end;
    walk(tree_of<W8>(s), r, bound, st, nodes_visited, [&](uint32_t ref) {
        uint32_t first, count;
        leaf_range(ref, &first, &count);
        for (uint32_t k = 0; k < count; ++k) {
            const BvhTriangle tri = s.tris[first + k];
            if (!(__float_as_uint(tri.v2.w) & TRI_NON_OPAQUE)) continue;
            const uint32_t inst = __float_as_uint(tri.v0.w), prim = __float_as_uint(tri.v1.w);
            if (inst == target_inst && prim == target_prim) continue;
            float t, b1, b2;
            if (!hit_triangle(w, xyz(tri.v0), xyz(tri.v1), xyz(tri.v2), t_end, &t, &b1, &b2)) continue;
            if (!(t < t_end)) continue;
            log_sum += int32_t(s.instances[inst].log2_transmittance);
            if (log_sum < ATTENUATION_CUTOFF) return true;
        }
        return false;
    });
    return log_sum < ATTENUATION_CUTOFF ? 0.0f : exp2_fixed(log_sum);
}

}  // namespace cn
