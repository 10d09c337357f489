// cn_math.h — float3 arithmetic and the Sobol-Burley sampler for the
// gfx950 stage kernels.
//
// Floating-point conventions: the library is compiled with
// -ffp-contract=off, so every operation below rounds where it is written;
// fmaf appears only where the reference kernels call fma. dot products
// accumulate x, y, z left to right; normalize(v) = v * (1/sqrt(v.v));
// lerp(x, y, a) = x*(1-a) + y*a (the SPIR-V FMix definition the reference's
// `lerp` lowers to). Division and sqrt are IEEE correctly rounded (HIP's
// default for f32).
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

namespace cn {

struct v3 {
    float x, y, z;
};

__host__ __device__ __forceinline__ v3 V3(float x, float y, float z) { return {x, y, z}; }
__host__ __device__ __forceinline__ v3 S3(float s) { return {s, s, s}; }
__host__ __device__ __forceinline__ v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
__host__ __device__ __forceinline__ v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__host__ __device__ __forceinline__ v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ __forceinline__ v3 operator*(float s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ __forceinline__ v3 operator/(v3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__host__ __device__ __forceinline__ v3 operator/(v3 a, v3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
__host__ __device__ __forceinline__ v3 operator/(float s, v3 b) { return {s / b.x, s / b.y, s / b.z}; }
__host__ __device__ __forceinline__ v3 operator+(v3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
__host__ __device__ __forceinline__ v3 operator+(float s, v3 a) { return {s + a.x, s + a.y, s + a.z}; }
__host__ __device__ __forceinline__ v3 operator-(float s, v3 a) { return {s - a.x, s - a.y, s - a.z}; }
__host__ __device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ v3 cross(v3 a, v3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ __forceinline__ float length(v3 v) { return sqrtf(dot(v, v)); }
__host__ __device__ __forceinline__ v3 normalize(v3 v) { return v * (1.0f / sqrtf(dot(v, v))); }
__host__ __device__ __forceinline__ float rsqrt_exact(float x) { return 1.0f / sqrtf(x); }
__host__ __device__ __forceinline__ v3 vabs(v3 v) { return {fabsf(v.x), fabsf(v.y), fabsf(v.z)}; }
__host__ __device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__host__ __device__ __forceinline__ float saturate(float x) { return clampf(x, 0.0f, 1.0f); }
__host__ __device__ __forceinline__ v3 vclamp(v3 v, float lo, float hi) {
    return {clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi)};
}
__host__ __device__ __forceinline__ v3 vmax(v3 a, v3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
__host__ __device__ __forceinline__ float lerpf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__host__ __device__ __forceinline__ v3 vlerp(v3 x, v3 y, float a) {
    const float b = 1.0f - a;
    return {x.x * b + y.x * a, x.y * b + y.y * a, x.z * b + y.z * a};
}
__host__ __device__ __forceinline__ bool any_positive(v3 v) { return v.x > 0.0f || v.y > 0.0f || v.z > 0.0f; }
__host__ __device__ __forceinline__ v3 vfma(v3 a, v3 b, v3 c) {
    return {fmaf(a.x, b.x, c.x), fmaf(a.y, b.y, c.y), fmaf(a.z, b.z, c.z)};
}
__host__ __device__ __forceinline__ float mean3(v3 v) { return (v.x + v.y + v.z) / 3.0f; }
__host__ __device__ __forceinline__ float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
__host__ __device__ __forceinline__ v3 xyz(float4 v) { return {v.x, v.y, v.z}; }
__host__ __device__ __forceinline__ float4 F4(v3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// --- Sobol-Burley (shaders/rng.slang) ---------------------------------------

// rng.slang:29-47 — the dimension registry.
constexpr uint32_t DIM_CAMERA_JITTER = 0;
constexpr uint32_t DIM_CAMERA_LENS = 1;
constexpr uint32_t SLOT_BSDF_LOBE = 0;
constexpr uint32_t SLOT_BSDF_DIR = 1;
constexpr uint32_t SLOT_NEE_LIGHT = 2;
constexpr uint32_t SLOT_NEE_POINT = 3;
constexpr uint32_t SLOT_RR = 4;
__host__ __device__ __forceinline__ uint32_t bounce_dim(uint32_t bounce, uint32_t slot) {
    return 8u + bounce * 8u + slot;
}

__host__ __device__ __forceinline__ uint32_t bit_reverse(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bitreverse32(x);
#else
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    return (x >> 16) | (x << 16);
#endif
}

// rng.slang:68-75
__host__ __device__ __forceinline__ uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x21f0aaadu;
    x ^= x >> 15;
    x *= 0xd35a2d97u;
    x ^= x >> 15;
    return x;
}

// rng.slang:81-95: Burley's Laine-Karras permutation applied in reversed
// bit order — the nested uniform (Owen) scramble.
__host__ __device__ __forceinline__ uint32_t owen_scramble(uint32_t x, uint32_t seed) {
    x = bit_reverse(x);
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return bit_reverse(x);
}

// rng.slang:100-111, axis 1: the generator matrix is Pascal's triangle mod
// 2. Its column for index bit k is row k's binomial parities from the top
// bit down, the coefficients of (1 + z)^k with z^j at bit 31 - j, so the
// value is the index polynomial f(t) = sum of t^k over its set bits taken
// at t = 1 + z, over GF(2). Since (1 + z)^(2^m) = 1 + z^(2^m) there,
// splitting f into halves gives f(1 + z) = (f_lo + f_hi)(1 + z) + z^h
// f_hi(1 + z): five steps of "each block's low half ^= its high half", at
// block sizes 2, 4, ..., 32, then a bit reversal for the top-down order —
// equal to the reference's bit-by-bit column loop for every 32-bit index
// (checked exhaustively), in 16 operations instead of up to 32 dependent
// loop trips.
__host__ __device__ __forceinline__ uint32_t sobol_axis1(uint32_t index) {
    uint32_t w = index;
    w ^= (w >> 1) & 0x55555555u;
    w ^= (w >> 2) & 0x33333333u;
    w ^= (w >> 4) & 0x0f0f0f0fu;
    w ^= (w >> 8) & 0x00ff00ffu;
    w ^= (w >> 16) & 0x0000ffffu;
    return bit_reverse(w);
}

// rng.slang:114-116
__host__ __device__ __forceinline__ float unit_float(uint32_t x) {
    return float(x >> 8) * (1.0f / 16777216.0f);
}

// rng.slang:128-132
__host__ __device__ __forceinline__ float hash_uniform(uint32_t pixel, uint32_t sample,
                                                       uint32_t bounce, uint32_t crossing) {
    uint32_t h = lowbias32(pixel ^ lowbias32(sample ^ 0x9e3779b9u));
    h = lowbias32(h ^ (bounce << 8 | crossing));
    return unit_float(h);
}

// Stochastic opacity's Bernoulli trial for one crossing of a
// fractional-opacity instance. The reference (intersect.slang:64-76) keys
// its draw on the crossing's rank in the driver's traversal order, which no
// other traversal reproduces. Here the draw is the hashUniform stream of
// (pixel, sample, bounce) with the crossing's (instance, primitive) folded
// in where the rank went: every surface a ray meets makes one independent
// trial whatever order the BVH is walked in, so the hit is the nearest
// crossing whose trial commits — the distribution the reference's comment
// describes ("the nearest crossing that commits wins").
__host__ __device__ __forceinline__ float crossing_uniform(uint32_t pixel, uint32_t sample, uint32_t bounce,
                                                           uint32_t instance, uint32_t primitive) {
    uint32_t h = lowbias32(pixel ^ lowbias32(sample ^ 0x9e3779b9u));
    h = lowbias32(h ^ (bounce << 8));
    h = lowbias32(h ^ lowbias32(instance ^ lowbias32(primitive ^ 0x85ebca6bu)));
    return unit_float(h);
}

// Fractional-opacity shadows multiply (1 - opacity) per crossing
// (trace_shadow.slang:52-65). A float product depends on the order of its
// factors and traversals meet crossings in different orders, so the product
// is kept as a sum of base-2 logarithms in fixed point (2^-24 units): integer
// adds, equal in any order. Each instance's log2(1 - opacity) is rounded to
// that grid once, on the host (log2_transmittance_fixed, cn_prep.cpp), and
// exp2_fixed maps a sum back. The reference ends a connection once its
// attenuation falls below 1e-4; here once the sum falls below
// ATTENUATION_CUTOFF = ceil(log2(1e-4) * 2^24). The sum never rises, so a
// walk may stop as soon as it is below.
constexpr int32_t LOG2_FIXED_SHIFT = 24;
constexpr int32_t LOG2_FIXED_FLOOR = -(1 << 28);  // one crossing of transmittance <= 2^-16
constexpr int32_t ATTENUATION_CUTOFF = -222930820;

// 2^(L / 2^24) for LOG2_FIXED_FLOOR * 2 <= L <= 0: e^x on [0, ln 2) by its
// Taylor polynomial to x^12 (truncation below 1e-11 relative) in double,
// Horner form, rounded once to float, times an exact power of two. Only
// IEEE adds and multiplies: the same bits on every processor built without
// contraction.
__host__ __device__ inline float exp2_fixed(int32_t L) {
    const int32_t n = L >> LOG2_FIXED_SHIFT;  // floor(L / 2^24)
    const int32_t f = L - n * (1 << LOG2_FIXED_SHIFT);
    const double x = double(f) * (0.6931471805599453 / 16777216.0);
    double p = 2.08767569878681e-09;  // 1/12!
    p = p * x + 2.505210838544172e-08;
    p = p * x + 2.755731922398589e-07;
    p = p * x + 2.7557319223985893e-06;
    p = p * x + 2.48015873015873e-05;
    p = p * x + 0.0001984126984126984;
    p = p * x + 0.001388888888888889;
    p = p * x + 0.008333333333333333;
    p = p * x + 0.041666666666666664;
    p = p * x + 0.16666666666666666;
    p = p * x + 0.5;
    p = p * x + 1.0;
    p = p * x + 1.0;
    union {
        uint32_t u;
        float f;
    } scale;
    scale.u = uint32_t(127 + (n < -126 ? -126 : n)) << 23;  // 2^n, exact
    return float(p) * scale.f;
}

struct f2 {
    float x, y;
};

// rng.slang:140-145
__host__ __device__ __forceinline__ f2 sample_2d(uint32_t pixel, uint32_t sample, uint32_t dimension) {
    const uint32_t seed = lowbias32(pixel ^ lowbias32(dimension));
    const uint32_t index = owen_scramble(sample, lowbias32(seed ^ 0xf8ade99au));
    const float a = unit_float(owen_scramble(bit_reverse(index), lowbias32(seed ^ 0xe0aaaf76u)));
    const float b = unit_float(owen_scramble(sobol_axis1(index), lowbias32(seed ^ 0x94964d4eu)));
    return {a, b};
}

// rng.slang:148-152
__host__ __device__ __forceinline__ float sample_1d(uint32_t pixel, uint32_t sample, uint32_t dimension) {
    const uint32_t seed = lowbias32(pixel ^ lowbias32(dimension));
    const uint32_t index = owen_scramble(sample, lowbias32(seed ^ 0xbff95bfeu));
    return unit_float(owen_scramble(bit_reverse(index), lowbias32(seed ^ 0x635fc02fu)));
}

}  // namespace cn
