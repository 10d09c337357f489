// cn_oracle.cpp — TEST INFRASTRUCTURE ONLY: the CPU oracle.
//
// A scalar, line-by-line restatement of the reference estimator and its host
// table builders, used by tests/ (and bench.py's cpu_baseline leg) to check
// the HIP path. Each function cites the reference file:line it restates;
// paths are relative to the reference repository root.
//
// Floating-point conventions (shared with the HIP path, so per-sample
// results can be compared bit for bit wherever no transcendental differs):
//   * no contraction: built with -ffp-contract=off; fmaf appears exactly
//     where the reference writes fma;
//   * dot(a, b) = (a.x*b.x + a.y*b.y) + a.z*b.z; normalize(v) =
//     v * (1/sqrt(dot(v, v))); rsqrt(x) = 1/sqrt(x); lerp(x, y, a) =
//     x*(1-a) + y*a (SPIR-V FMix);
//   * ray/triangle intersection is the watertight test of Woop, Benthin &
//     Wald (JCGT 2013), closest hit, ties broken toward the lower
//     (instance, primitive) — the reference delegates this to the driver.
//
// Parity pins for each part are listed in the test files that load this
// library (tests/test_oracle_*.py).

#include "cn_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- vectors

struct f3 {
    float x, y, z;
};
struct f2 {
    float x, y;
};
struct f4 {
    float x, y, z, w;
};

inline f3 V(float x, float y, float z) { return {x, y, z}; }
inline f3 splat(float s) { return {s, s, s}; }
inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
inline f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline f3 operator/(f3 a, f3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline f3 operator/(float s, f3 b) { return {s / b.x, s / b.y, s / b.z}; }
inline f3 operator+(f3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
inline f3 operator+(float s, f3 a) { return {s + a.x, s + a.y, s + a.z}; }
inline f3 operator-(float s, f3 a) { return {s - a.x, s - a.y, s - a.z}; }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(f3 v) { return std::sqrt(dot(v, v)); }
inline f3 normalize(f3 v) { return v * (1.0f / std::sqrt(dot(v, v))); }
inline float rsqrt(float x) { return 1.0f / std::sqrt(x); }
inline f3 abs3(f3 v) { return {std::fabs(v.x), std::fabs(v.y), std::fabs(v.z)}; }
inline float clampf(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }
inline float saturate(float x) { return clampf(x, 0.0f, 1.0f); }
inline f3 clamp3(f3 v, float lo, float hi) {
    return {clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi)};
}
inline f3 max3(f3 a, f3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
inline float lerp(float x, float y, float a) { return x * (1.0f - a) + y * a; }
inline f3 lerp3(f3 x, f3 y, float a) {
    float b = 1.0f - a;
    return {x.x * b + y.x * a, x.y * b + y.y * a, x.z * b + y.z * a};
}
inline bool any_pos(f3 v) { return v.x > 0.0f || v.y > 0.0f || v.z > 0.0f; }
inline f3 fma3(f3 a, f3 b, f3 c) {
    return {std::fmaf(a.x, b.x, c.x), std::fmaf(a.y, b.y, c.y), std::fmaf(a.z, b.z, c.z)};
}
inline f3 exp3(f3 v) { return {std::exp(v.x), std::exp(v.y), std::exp(v.z)}; }
inline f3 log3(f3 v) { return {std::log(v.x), std::log(v.y), std::log(v.z)}; }
inline float average(f3 v) { return (v.x + v.y + v.z) / 3.0f; }
inline float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

constexpr float PI_SLANG = 3.14159265358979323846f;  // raygen.slang:12, environment.slang:25
constexpr float PI_OPENPBR = 3.14159265358979f;      // openpbr.slang:40

// ------------------------------------------------------------ rng.slang

// shaders/rng.slang:29-47
constexpr uint32_t CAMERA_JITTER = 0;
constexpr uint32_t CAMERA_LENS = 1;
constexpr uint32_t PATH_BASE = 8;
constexpr uint32_t BSDF_LOBE = 0;
constexpr uint32_t BSDF_DIR = 1;
constexpr uint32_t NEE_LIGHT = 2;
constexpr uint32_t NEE_POINT = 3;
constexpr uint32_t RR = 4;
constexpr uint32_t BOUNCE_SLOTS = 8;
inline uint32_t bounce_dimension(uint32_t bounce, uint32_t slot) {
    return PATH_BASE + bounce * BOUNCE_SLOTS + slot;
}

// shaders/rng.slang:51-64: the generator matrix of the Sobol sequence's
// second dimension is Pascal's triangle mod 2. Derived here from Lucas'
// theorem (C(k, j) is odd iff j & ~k == 0) rather than typed in, so a
// test comparing it with the reference's literal table pins both.
struct SobolDim1 {
    uint32_t column[32];
    SobolDim1() {
        for (uint32_t k = 0; k < 32; ++k) {
            uint32_t value = 0;
            for (uint32_t j = 0; j <= k; ++j) {
                if ((j & ~k) == 0) value |= 0x80000000u >> j;
            }
            column[k] = value;
        }
    }
};
const SobolDim1 SOBOL_DIM1;

inline uint32_t reverse_bits(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    return (x >> 16) | (x << 16);
}

// shaders/rng.slang:68-75
inline uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x21f0aaadu;
    x ^= x >> 15;
    x *= 0xd35a2d97u;
    x ^= x >> 15;
    return x;
}

// shaders/rng.slang:81-88
inline uint32_t laine_karras(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}

// shaders/rng.slang:93-95
inline uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
    return reverse_bits(laine_karras(reverse_bits(x), seed));
}

// shaders/rng.slang:100-111
inline uint32_t sobol(uint32_t index, uint32_t axis) {
    if (axis == 0) return reverse_bits(index);
    uint32_t value = 0;
    for (uint32_t bit = 0; index != 0; ++bit, index >>= 1) {
        if (index & 1u) value ^= SOBOL_DIM1.column[bit];
    }
    return value;
}

// shaders/rng.slang:114-116
inline float to_unit_float(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }

// shaders/rng.slang:119-121
inline float scrambled_sobol(uint32_t index, uint32_t axis, uint32_t seed) {
    return to_unit_float(nested_uniform_scramble(sobol(index, axis), seed));
}

// shaders/rng.slang:128-132
inline float hash_uniform(uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t crossing) {
    uint32_t h = lowbias32(pixel ^ lowbias32(sample ^ 0x9e3779b9u));
    h = lowbias32(h ^ (bounce << 8 | crossing));
    return to_unit_float(h);
}

// Stochastic opacity's per-crossing Bernoulli trial. intersect.slang:64-76
// draws hashUniform(pixel, sample, bounce, crossing) with `crossing` the
// candidate's rank in the driver's traversal order; that rank is not
// reproducible outside the driver, so the build keys the trial on the
// crossing's identity instead: the same (pixel, sample, bounce) hash chain
// with (instance, primitive) folded in where the rank went. Every surface
// then makes one independent trial and the hit is the nearest surface whose
// trial commits, in any traversal order.
inline float crossing_uniform(uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t instance,
                              uint32_t primitive) {
    uint32_t h = lowbias32(pixel ^ lowbias32(sample ^ 0x9e3779b9u));
    h = lowbias32(h ^ (bounce << 8));
    h = lowbias32(h ^ lowbias32(instance ^ lowbias32(primitive ^ 0x85ebca6bu)));
    return to_unit_float(h);
}

// trace_shadow.slang:52-65 multiplies (1 - opacity) per crossing and gives
// up below 1e-4. The float product's value depends on the order of its
// factors, so the build keeps it as a sum of base-2 logarithms in 2^-24
// fixed point (integer adds, the same in any order): per instance
// log2(1 - opacity) rounded to nearest on that grid, floored at -2^28; the
// connection is dark once the sum is below ceil(log2(1e-4) * 2^24); a sum
// maps back to a float by a fixed Taylor polynomial of e^x in double.
constexpr int32_t LOG2_FIXED_SHIFT = 24;
constexpr int32_t LOG2_FIXED_FLOOR = -(1 << 28);
constexpr int32_t ATTENUATION_CUTOFF = -222930820;

int32_t log2_transmittance_fixed(float opacity) {
    double transmittance = 1.0 - double(opacity);
    if (!(transmittance > 0.0)) return LOG2_FIXED_FLOOR;
    if (transmittance >= 1.0) return 0;
    double units = std::nearbyint(std::log2(transmittance) * double(1 << LOG2_FIXED_SHIFT));
    return units < double(LOG2_FIXED_FLOOR) ? LOG2_FIXED_FLOOR : int32_t(units);
}

float exp2_fixed(int32_t L) {
    int32_t n = L >> LOG2_FIXED_SHIFT;
    int32_t f = L - n * (1 << LOG2_FIXED_SHIFT);
    double x = double(f) * (0.6931471805599453 / 16777216.0);
    // 1/k! for k = 12 down to 0
    static const double c[13] = {2.08767569878681e-09, 2.505210838544172e-08, 2.755731922398589e-07,
                                 2.7557319223985893e-06, 2.48015873015873e-05, 0.0001984126984126984,
                                 0.001388888888888889, 0.008333333333333333, 0.041666666666666664,
                                 0.16666666666666666, 0.5, 1.0, 1.0};
    double p = c[0];
    for (int k = 1; k < 13; ++k) p = p * x + c[k];
    return std::ldexp(float(p), n < -126 ? -126 : n);
}

// shaders/rng.slang:140-145
inline f2 sample_2d(uint32_t pixel, uint32_t sample, uint32_t dimension) {
    uint32_t seed = lowbias32(pixel ^ lowbias32(dimension));
    uint32_t index = nested_uniform_scramble(sample, lowbias32(seed ^ 0xf8ade99au));
    return {scrambled_sobol(index, 0, lowbias32(seed ^ 0xe0aaaf76u)),
            scrambled_sobol(index, 1, lowbias32(seed ^ 0x94964d4eu))};
}

// shaders/rng.slang:148-152
inline float sample_1d(uint32_t pixel, uint32_t sample, uint32_t dimension) {
    uint32_t seed = lowbias32(pixel ^ lowbias32(dimension));
    uint32_t index = nested_uniform_scramble(sample, lowbias32(seed ^ 0xbff95bfeu));
    return scrambled_sobol(index, 0, lowbias32(seed ^ 0x635fc02fu));
}

// ----------------------------------------------------------- color.rs

// crates/cenote/src/color.rs:29-31 — glam Vec3::dot, left to right.
inline float luminance(f3 c) { return c.x * 0.2722287f + c.y * 0.6740818f + c.z * 0.0536895f; }

// color.rs:13-17, as glam's Mat3 * Vec3: x_axis*x + y_axis*y + z_axis*z.
inline f3 acescg_from_rec709(f3 c) {
    const f3 cx = V(0.6130974f, 0.07019373f, 0.020615594f);
    const f3 cy = V(0.33952314f, 0.9163539f, 0.10956978f);
    const f3 cz = V(0.04737945f, 0.013452399f, 0.86981466f);
    return (cx * c.x + cy * c.y) + cz * c.z;
}

// --------------------------------------------------------- lights.rs

constexpr uint32_t LIGHT_NONE = 0xffffffffu;  // lights.rs:28
constexpr uint32_t KIND_TRIANGLE = 0;         // lights.rs:32-34
constexpr uint32_t KIND_DISTANT = 1;
constexpr uint32_t KIND_POINT = 2;

struct LightRec {
    f3 anchor;
    float pdf;
    f3 edge1;
    float alias_threshold;
    f3 edge2;
    uint32_t alias_index;
    f3 emission;
    uint32_t instance;
    uint32_t kind;
    uint32_t primitive;
    uint32_t pad0, pad1;
};
static_assert(sizeof(LightRec) == 80, "lights.rs:115");

// lights.rs:135-137: glam cross (scalar Vec3) and length = sqrt(dot).
inline double triangle_area(const LightRec& r) {
    return double(length(cross(r.edge1, r.edge2))) / 2.0;
}

// lights.rs:123-130
inline double light_power(const LightRec& r) {
    double luma = double(luminance(r.emission));
    switch (r.kind) {
        case KIND_TRIANGLE: return M_PI * luma * triangle_area(r);
        case KIND_DISTANT: return luma;
        default: return 4.0 * M_PI * luma;
    }
}

struct TriangleLight {
    f3 corners[3];
    f3 emission;
    uint32_t instance, primitive;
};

// lights.rs:141-162 and :236-275
std::vector<LightRec> raw_records(const std::vector<TriangleLight>& tris,
                                  const cn_delta_light* deltas, uint32_t delta_count) {
    std::vector<LightRec> records;
    records.reserve(tris.size() + delta_count);
    auto record = [](uint32_t kind, f3 anchor, f3 e1, f3 e2, f3 emission, uint32_t inst,
                     uint32_t prim) {
        LightRec r{};
        r.anchor = anchor;
        r.pdf = 0.0f;
        r.edge1 = e1;
        r.alias_threshold = 1.0f;
        r.edge2 = e2;
        r.alias_index = 0;
        r.emission = emission;
        r.instance = inst;
        r.kind = kind;
        r.primitive = prim;
        return r;
    };
    for (const auto& t : tris) {
        records.push_back(record(KIND_TRIANGLE, t.corners[0], t.corners[1] - t.corners[0],
                                 t.corners[2] - t.corners[0], t.emission, t.instance,
                                 t.primitive));
    }
    for (uint32_t i = 0; i < delta_count; ++i) {
        const cn_delta_light& d = deltas[i];
        f3 v = V(d.vector[0], d.vector[1], d.vector[2]);
        f3 c = V(d.color[0], d.color[1], d.color[2]);
        records.push_back(record(d.kind == CN_LIGHT_DISTANT ? KIND_DISTANT : KIND_POINT, v,
                                 splat(0.0f), splat(0.0f), c, LIGHT_NONE, 0));
    }
    return records;
}

// lights.rs:175-224: power-proportional Walker/Vose alias table.
std::vector<LightRec> build_lights(const std::vector<TriangleLight>& tris,
                                   const cn_delta_light* deltas, uint32_t delta_count,
                                   double* total_out) {
    std::vector<LightRec> records = raw_records(tris, deltas, delta_count);
    std::vector<double> powers(records.size());
    for (size_t i = 0; i < records.size(); ++i) powers[i] = light_power(records[i]);
    double total = 0.0;  // Iterator::sum: a left fold
    for (double p : powers) total += p;
    if (total_out) *total_out = total;
    if (total <= 0.0) return records;
    for (size_t i = 0; i < records.size(); ++i) {
        double selection = powers[i] / total;
        if (records[i].kind == KIND_TRIANGLE) {
            double area = triangle_area(records[i]);
            records[i].pdf = area > 0.0 ? float(selection / area) : 0.0f;
        } else {
            records[i].pdf = float(selection);
        }
    }
    double n = double(powers.size());
    std::vector<double> share(powers.size());
    for (size_t i = 0; i < powers.size(); ++i) share[i] = powers[i] * n / total;
    // `partition` keeps index order in both halves; `pop` takes from the end.
    std::vector<size_t> small, large;
    for (size_t i = 0; i < share.size(); ++i) (share[i] < 1.0 ? small : large).push_back(i);
    while (!small.empty() && !large.empty()) {
        size_t poor = small.back();
        small.pop_back();
        size_t rich = large.back();
        large.pop_back();
        records[poor].alias_threshold = float(share[poor]);
        records[poor].alias_index = uint32_t(rich);
        share[rich] -= 1.0 - share[poor];
        (share[rich] < 1.0 ? small : large).push_back(rich);
    }
    return records;
}

// scene/mod.rs:665-671
inline float select_probability(double env_power, double light_power) {
    if (light_power == 0.0) return env_power > 0.0 ? 1.0f : 0.0f;
    return float(env_power / (env_power + light_power));
}

// ---------------------------------------------------- environment.rs

struct EnvTables {
    std::vector<float> marginal, conditional, pdfs;
    double power = 0.0;
};

// environment.rs:130-218
EnvTables env_tables(uint32_t width, uint32_t height, const float* texels) {
    const size_t w = width, h = height;
    std::vector<double> lum(w * h);
    for (size_t i = 0; i < w * h; ++i) {
        lum[i] = double(luminance(V(texels[4 * i], texels[4 * i + 1], texels[4 * i + 2])));
    }
    std::vector<double> weight(w * h, 0.0);
    for (size_t row = 0; row < h; ++row) {
        double sin_theta = std::sin((double(row) + 0.5) / double(h) * M_PI);
        for (size_t col = 0; col < w; ++col) {
            double peak = 0.0;
            for (int dr = -1; dr <= 1; ++dr) {
                // saturating_add_signed then min(h - 1)
                long rr = long(row) + dr;
                if (rr < 0) rr = 0;
                size_t r = std::min(size_t(rr), h - 1);
                for (int dc = -1; dc <= 1; ++dc) {
                    long cc = (long(col) + dc) % long(w);
                    if (cc < 0) cc += long(w);
                    peak = std::max(peak, lum[r * w + size_t(cc)]);
                }
            }
            weight[row * w + col] = peak * sin_theta;
        }
    }
    double total = 0.0;
    for (double v : weight) total += v;

    EnvTables t;
    t.marginal.assign(h + 1, 0.0f);
    double running = 0.0;
    for (size_t row = 0; row < h; ++row) {
        double row_sum = 0.0;
        for (size_t col = 0; col < w; ++col) row_sum += weight[row * w + col];
        running += row_sum;
        t.marginal[row + 1] = total > 0.0 ? float(running / total) : 0.0f;
    }
    t.marginal[h] = 1.0f;

    t.conditional.assign(h * (w + 1), 0.0f);
    for (size_t row = 0; row < h; ++row) {
        double row_sum = 0.0;
        for (size_t col = 0; col < w; ++col) row_sum += weight[row * w + col];
        float* cdf = &t.conditional[row * (w + 1)];
        double run = 0.0;
        for (size_t col = 0; col < w; ++col) {
            run += weight[row * w + col];
            cdf[col + 1] = row_sum > 0.0 ? float(run / row_sum) : 0.0f;
        }
        cdf[w] = 1.0f;
    }

    double scale = total > 0.0 ? double(w * h) / total : 0.0;
    t.pdfs.resize(w * h);
    for (size_t i = 0; i < w * h; ++i) t.pdfs[i] = float(weight[i] * scale);

    double power = 0.0;
    for (size_t row = 0; row < h; ++row) {
        double theta0 = double(row) / double(h) * M_PI;
        double theta1 = double(row + 1) / double(h) * M_PI;
        double solid = 2.0 * M_PI / double(w) * (std::cos(theta0) - std::cos(theta1));
        double s = 0.0;
        for (size_t col = 0; col < w; ++col) s += lum[row * w + col];
        power += solid * s;
    }
    t.power = power;
    return t;
}

// ------------------------------------------------------------ camera

struct Basis {
    f3 right, up, forward;
};

// scene/mod.rs:448-461 with glam's normalize (v * (1/length)) and the
// lens pre-scaling of wavefront.rs:643-653; focus distance from
// scene/lower.rs:227-232.
Basis camera_basis(const cn_camera& c, float aspect, float* aperture) {
    f3 position = V(c.position[0], c.position[1], c.position[2]);
    f3 look_at = V(c.look_at[0], c.look_at[1], c.look_at[2]);
    f3 upv = V(c.up[0], c.up[1], c.up[2]);
    f3 forward = normalize(look_at - position);
    f3 right = normalize(cross(forward, upv));
    f3 up = cross(right, forward);
    const float rads = 3.14159265358979323846f / 180.0f;  // f32::to_radians
    float half_height = std::tan(c.vfov_degrees * rads / 2.0f);
    Basis b{(right * half_height) * aspect, up * half_height, forward};
    *aperture = 0.0f;
    if (c.aperture_radius > 0.0f) {
        float focus = c.focus_distance > 0.0f ? c.focus_distance : length(position - look_at);
        b.right = b.right * focus;
        b.up = b.up * focus;
        b.forward = b.forward * focus;
        *aperture = c.aperture_radius;
    }
    return b;
}

// -------------------------------------------------------- openpbr.slang

struct Material {
    f3 baseColor;
    float baseRoughness;
    f3 emission;
    float metalness;
    float specularWeight;
    float specularRoughness;
    float specularIor;
    float transmissionWeight;
    f3 transmissionColor;
    float transmissionDepth;
    f3 coatColor;
    float coatWeight;
    float coatRoughness;
    float coatIor;
    float coatDarkening;
    float fuzzWeight;
    f3 fuzzColor;
    float fuzzRoughness;
    float opacity;
    uint32_t thinWalled;
};

Material lift_material(const cn_material& m) {
    Material r;
    r.baseColor = V(m.base_color[0], m.base_color[1], m.base_color[2]);
    r.baseRoughness = m.base_roughness;
    r.emission = V(m.emission[0], m.emission[1], m.emission[2]);
    r.metalness = m.metalness;
    r.specularWeight = m.specular_weight;
    r.specularRoughness = m.specular_roughness;
    r.specularIor = m.specular_ior;
    r.transmissionWeight = m.transmission_weight;
    r.transmissionColor =
        V(m.transmission_color[0], m.transmission_color[1], m.transmission_color[2]);
    r.transmissionDepth = m.transmission_depth;
    r.coatColor = V(m.coat_color[0], m.coat_color[1], m.coat_color[2]);
    r.coatWeight = m.coat_weight;
    r.coatRoughness = m.coat_roughness;
    r.coatIor = m.coat_ior;
    r.coatDarkening = m.coat_darkening;
    r.fuzzWeight = m.fuzz_weight;
    r.fuzzColor = V(m.fuzz_color[0], m.fuzz_color[1], m.fuzz_color[2]);
    r.fuzzRoughness = m.fuzz_roughness;
    r.opacity = m.opacity;
    r.thinWalled = m.thin_walled;
    return r;
}

constexpr float MIN_SPECULAR_ROUGHNESS = 0.035f;  // openpbr.slang:45

// openpbr.slang:86-95
constexpr uint32_t TABLE_GGX_E = 0;
constexpr uint32_t TABLE_GGX_E_AVG = 1024;
constexpr uint32_t TABLE_GLASS_E = 1056;
constexpr uint32_t TABLE_GLASS_E_AVG = 5152;
constexpr uint32_t TABLE_GLASS_INV_E = 5408;
constexpr uint32_t TABLE_GLASS_INV_E_AVG = 9504;
constexpr uint32_t TABLE_DIELECTRIC_E = 9760;
constexpr uint32_t TABLE_SHEEN_A = 13856;
constexpr uint32_t TABLE_SHEEN_B = 14880;
constexpr uint32_t TABLE_SHEEN_R = 15904;

// openpbr.slang:97-102
float tableRead1D(const float* t, uint32_t offset, uint32_t size, float x) {
    float fx = saturate(x) * float(size - 1);
    uint32_t i = std::min(uint32_t(fx), size - 1);
    uint32_t j = std::min(i + 1, size - 1);
    return lerp(t[offset + i], t[offset + j], fx - float(i));
}

// openpbr.slang:104-111
float tableRead2D(const float* t, uint32_t offset, uint32_t sx, uint32_t sy, float x, float y) {
    float fy = saturate(y) * float(sy - 1);
    uint32_t i = std::min(uint32_t(fy), sy - 1);
    uint32_t j = std::min(i + 1, sy - 1);
    return lerp(tableRead1D(t, offset + sx * i, sx, x), tableRead1D(t, offset + sx * j, sx, x),
                fy - float(i));
}

// openpbr.slang:113-121
float tableRead3D(const float* t, uint32_t offset, uint32_t sx, uint32_t sy, uint32_t sz, float x,
                  float y, float z) {
    float fz = saturate(z) * float(sz - 1);
    uint32_t i = std::min(uint32_t(fz), sz - 1);
    uint32_t j = std::min(i + 1, sz - 1);
    uint32_t plane = sx * sy;
    return lerp(tableRead2D(t, offset + plane * i, sx, sy, x, y),
                tableRead2D(t, offset + plane * j, sx, sy, x, y), fz - float(i));
}

// openpbr.slang:125-127
inline float iorRemap(float eta) { return std::sqrt(std::fabs(eta - 1.0f) / (eta + 1.0f)); }

// openpbr.slang:134-135
const float FON_A = 0.5f - 2.0f / (3.0f * PI_OPENPBR);
const float FON_B = 2.0f / 3.0f - 28.0f / (15.0f * PI_OPENPBR);

// openpbr.slang:140-145
inline float fonAlbedo(float mu, float r) {
    float g = 1.0f - mu;
    float gOverPi =
        g * (0.0571085289f + g * (0.491881867f + g * (-0.332181442f + g * 0.0714429953f)));
    return (1.0f + r * gOverPi) / (1.0f + FON_A * r);
}

// openpbr.slang:152-166
f3 eonBrdf(f3 rho, float r, f3 wi, f3 wo) {
    float s = dot(wi, wo) - wi.z * wo.z;
    float sOverT = s > 0.0f ? s / std::fmax(wi.z, wo.z) : s;
    float A = 1.0f / (1.0f + FON_A * r);
    f3 singleScatter = ((rho / PI_OPENPBR) * A) * (1.0f + r * sOverT);
    float albedoI = fonAlbedo(wi.z, r);
    float albedoO = fonAlbedo(wo.z, r);
    float albedoAvg = A * (1.0f + FON_B * r);
    f3 rhoMulti = ((rho * rho) * albedoAvg) / (splat(1.0f) - rho * (1.0f - albedoAvg));
    const float eps = 1e-7f;
    f3 multiScatter = (((rhoMulti / PI_OPENPBR) * std::fmax(eps, 1.0f - albedoO)) *
                       std::fmax(eps, 1.0f - albedoI)) /
                      std::fmax(eps, 1.0f - albedoAvg);
    return singleScatter + multiScatter;
}

// openpbr.slang:170-175
inline f4 sampleCosineHemisphere(f2 u) {
    float r = std::sqrt(u.x);
    float phi = 2.0f * PI_OPENPBR * u.y;
    float z = std::sqrt(1.0f - u.x);
    return {r * std::cos(phi), r * std::sin(phi), z, z / PI_OPENPBR};
}

// openpbr.slang:179-183
inline float ggxD(float alpha, float hz) {
    float a2 = alpha * alpha;
    float denom = hz * hz * (a2 - 1.0f) + 1.0f;
    return hz > 0.0f ? a2 / (PI_OPENPBR * denom * denom) : 0.0f;
}

// openpbr.slang:189-192
inline float smithG1(float alpha, float vz) {
    float a2 = alpha * alpha;
    return 2.0f * std::fabs(vz) / (std::fabs(vz) + std::sqrt(vz * vz * (1.0f - a2) + a2));
}

// openpbr.slang:198-210
f3 sampleVndf(f3 wo, float alpha, f2 u) {
    f3 woStd = normalize(V(wo.x * alpha, wo.y * alpha, wo.z));
    float phi = 2.0f * PI_OPENPBR * u.x;
    float z = std::fmaf(1.0f - u.y, 1.0f + woStd.z, -woStd.z);
    float sinTheta = std::sqrt(clampf(1.0f - z * z, 0.0f, 1.0f));
    f3 c = V(sinTheta * std::cos(phi), sinTheta * std::sin(phi), z);
    f3 hStd = c + woStd;
    return normalize(V(hStd.x * alpha, hStd.y * alpha, hStd.z));
}

// openpbr.slang:214-216
inline float vndfReflectPdf(float alpha, f3 wo, f3 h) {
    return smithG1(alpha, wo.z) * ggxD(alpha, h.z) / (4.0f * wo.z);
}

// openpbr.slang:221-225
inline f3 fresnelSchlick(f3 f0, float cosine) {
    float m = 1.0f - std::fabs(cosine);
    float m2 = m * m;
    return f0 + (((1.0f - f0) * m2) * m2) * m;
}

// openpbr.slang:228-230
inline f3 fresnelSchlickAverage(f3 f0) { return lerp3(f0, splat(1.0f), 1.0f / 21.0f); }

// openpbr.slang:235-248
float fresnelDielectric(float eta, float cosine) {
    float mu = std::fabs(cosine);
    float sin2 = 1.0f - mu * mu;
    float cos2t = 1.0f - sin2 / (eta * eta);
    if (cos2t < 0.0f) return 1.0f;
    float t0 = std::sqrt(cos2t);
    float t1 = eta * t0;
    float t2 = eta * mu;
    float rs = (mu - t1) / (mu + t1);
    float rp = (t0 - t2) / (t0 + t2);
    return (rs * rs + rp * rp) / 2.0f;
}

// openpbr.slang:253-258
inline float fresnelDielectricAverage(float eta) {
    if (eta < 1.0f) return 0.997118f + eta * (0.1014f - eta * (0.965241f + eta * 0.130607f));
    return (eta - 1.0f) / (4.08567f + 1.00071f * eta);
}

// openpbr.slang:264-266
inline f3 fresnelMultiScatter(f3 fss, float eAvg) {
    return (fss * eAvg) / (splat(1.0f) - fss * (1.0f - eAvg));
}

// openpbr.slang:270-277
inline f4 refractDirection(f3 wo, f3 h, float eta) {
    float c = dot(wo, h);
    float cos2t = 1.0f - (1.0f - c * c) / (eta * eta);
    if (cos2t <= 0.0f) return {0, 0, 0, 0};
    f3 r = h * (c / eta - std::sqrt(cos2t)) - wo / eta;
    return {r.x, r.y, r.z, 1.0f};
}

// openpbr.slang:289-293
constexpr uint32_t LOBE_DIFFUSE = 0;
constexpr uint32_t LOBE_SPECULAR = 1;
constexpr uint32_t LOBE_TRANSMISSION = 2;
constexpr uint32_t LOBE_COAT = 3;
constexpr uint32_t LOBE_FUZZ = 4;

// openpbr.slang:295-340
struct Closure {
    Material m;
    const float* tables;
    float alpha, coatAlpha, eta;
    bool exiting, hasSpecular;
    float fuzzA, fuzzB, fuzzR, fuzzAtten;
    float coatE, coatScale;
    f3 baseWeight;
    float specE;
    f3 conductorScale;
    float dielectricScale;
    f3 glassScale, glassTint;
    f3 thinR, thinT;
    float thinAlpha, thinScaleR, thinScaleT, thinTShare;
    float wFuzz, wCoat, wSpec, wDiffuse, wTotal;
    f3 aovAlbedo;
};

// openpbr.slang:348-352
f3 reflectionScale(const float* t, float rough, float mu, f3 fss) {
    float e = std::fmax(1e-4f, tableRead2D(t, TABLE_GGX_E, 32, 32, rough, mu));
    float eAvg = tableRead1D(t, TABLE_GGX_E_AVG, 32, rough);
    return 1.0f + fresnelMultiScatter(fss, eAvg) * ((1.0f - e) / e);
}

// openpbr.slang:357-363
float dielectricAlbedo(const float* t, float rough, float mu, float eta) {
    if (eta < 1.0f) return fresnelDielectric(eta, mu);
    return tableRead3D(t, TABLE_DIELECTRIC_E, 16, 16, 16, rough, mu, iorRemap(eta));
}

// openpbr.slang:371-384
f3 coatDarkening(const Material& m, float rough, float mu, float baseFresnel) {
    float fssCoat = fresnelDielectricAverage(m.coatIor);
    float kRough = 1.0f - (1.0f - fssCoat) / (m.coatIor * m.coatIor);
    float kSmooth = fresnelDielectric(m.coatIor, mu);
    float roughDiffuse = lerp(1.0f, rough, saturate(m.specularWeight * baseFresnel));
    float baseRough = lerp(roughDiffuse, rough, m.metalness);
    float k = lerp(kSmooth, kRough, baseRough);
    f3 baseAlbedo = lerp3(m.baseColor, splat(1.0f), (1.0f - m.metalness) * m.transmissionWeight);
    f3 darkened = (1.0f - k) / max3(splat(1e-4f), 1.0f - baseAlbedo * k);
    return lerp3(splat(1.0f), darkened, m.coatWeight * m.coatDarkening);
}

// openpbr.slang:391-587
Closure resolveClosure(const float* t, const Material& m, f3 wo, bool exiting) {
    Closure c;
    c.m = m;
    c.tables = t;
    c.exiting = exiting;
    c.hasSpecular = exiting || m.metalness > 0.0f || m.transmissionWeight > 0.0f ||
                    m.specularWeight > 0.0f;
    float mu = wo.z;

    float rough = clampf(m.specularRoughness, MIN_SPECULAR_ROUGHNESS, 1.0f);
    float coatRough = clampf(m.coatRoughness, MIN_SPECULAR_ROUGHNESS, 1.0f);
    if (m.coatWeight > 0.0f && !exiting) {
        float variance = std::fmin(1.0f, rough * rough * rough * rough +
                                             2.0f * coatRough * coatRough * coatRough * coatRough);
        rough = lerp(rough, std::sqrt(std::sqrt(variance)), m.coatWeight);
    }
    c.alpha = rough * rough;
    c.coatAlpha = coatRough * coatRough;

    float etaS = std::fmax(m.specularIor, 1e-4f);
    if (m.coatWeight > 0.0f && !exiting) {
        float etaBc = m.coatIor > etaS ? m.coatIor / etaS : etaS / m.coatIor;
        etaS = lerp(etaS, etaBc, m.coatWeight);
    }
    float f0 = (etaS - 1.0f) * (etaS - 1.0f) / ((etaS + 1.0f) * (etaS + 1.0f));
    float weight = clampf(m.specularWeight, 0.0f, f0 > 1e-6f ? 1.0f / f0 : 0.0f);
    float epsilon = signf(etaS - 1.0f) * std::sqrt(weight * f0);
    float eta = (1.0f + epsilon) / std::fmax(1e-3f, 1.0f - epsilon);
    if (std::fabs(eta - 1.0f) < 1e-3f) eta = eta < 1.0f ? 1.0f - 1e-3f : 1.0f + 1e-3f;
    c.eta = exiting ? 1.0f / eta : eta;

    c.fuzzA = 0.0f;
    c.fuzzB = 0.0f;
    c.fuzzR = 0.0f;
    c.fuzzAtten = 1.0f;
    if (m.fuzzWeight > 0.0f && !exiting) {
        float fuzzRough = clampf(m.fuzzRoughness, 1e-3f, 1.0f);
        float a = tableRead2D(t, TABLE_SHEEN_A, 32, 32, mu, fuzzRough);
        float r = tableRead2D(t, TABLE_SHEEN_R, 32, 32, mu, fuzzRough);
        if (std::fabs(a) >= 1e-5f && r >= 1e-5f) {
            c.fuzzA = a;
            c.fuzzB = tableRead2D(t, TABLE_SHEEN_B, 32, 32, mu, fuzzRough);
            c.fuzzR = r;
            c.fuzzAtten = 1.0f - m.fuzzWeight * r;
        }
    }

    c.coatE = 0.0f;
    c.coatScale = 1.0f;
    f3 coatAtten = splat(1.0f);
    if (m.coatWeight > 0.0f && !exiting) {
        c.coatE = dielectricAlbedo(t, coatRough, mu, m.coatIor);
        float fssCoat = fresnelDielectricAverage(m.coatIor);
        c.coatScale = reflectionScale(t, coatRough, mu, splat(fssCoat)).x;
        coatAtten = lerp3(splat(1.0f), m.coatColor * (1.0f - c.coatE), m.coatWeight) *
                    coatDarkening(m, rough, mu, f0);
    }
    c.baseWeight = splat(c.fuzzAtten) * coatAtten;

    c.specE = c.hasSpecular ? dielectricAlbedo(t, rough, mu, c.eta) : 0.0f;
    c.conductorScale = reflectionScale(t, rough, mu, fresnelSchlickAverage(m.baseColor));
    c.dielectricScale = reflectionScale(t, rough, mu, splat(fresnelDielectricAverage(c.eta))).x;

    c.glassTint = m.transmissionDepth == 0.0f ? m.transmissionColor : splat(1.0f);
    float glassEta = c.eta;
    uint32_t glassOffset = TABLE_GLASS_E;
    uint32_t glassAvgOffset = TABLE_GLASS_E_AVG;
    if (glassEta < 1.0f) {
        glassEta = 1.0f / glassEta;
        glassOffset = TABLE_GLASS_INV_E;
        glassAvgOffset = TABLE_GLASS_INV_E_AVG;
    }
    float glassZ = iorRemap(glassEta);
    float glassE = std::fmax(1e-4f, tableRead3D(t, glassOffset, 16, 16, 16, rough, mu, glassZ));
    float glassEAvg = tableRead2D(t, glassAvgOffset, 16, 16, rough, glassZ);
    c.glassScale = 1.0f + fresnelMultiScatter(c.glassTint, glassEAvg) * ((1.0f - glassE) / glassE);

    c.thinR = splat(0.0f);
    c.thinT = splat(0.0f);
    c.thinAlpha = c.alpha;
    c.thinScaleR = 1.0f;
    c.thinScaleT = 1.0f;
    c.thinTShare = 0.0f;
    bool thin = m.thinWalled != 0 && m.transmissionWeight > 0.0f;
    if (thin) {
        float thinEta = std::fmax(eta, 1.0f / eta);
        float r1 = fresnelDielectric(thinEta, mu);
        float t1 = 1.0f - r1;
        float cos2t = std::fmax(0.0f, 1.0f - (1.0f - mu * mu) / (thinEta * thinEta));
        f3 absorb = splat(0.0f);
        if (cos2t > 0.0f) {
            f3 tc = clamp3(m.transmissionColor, 1e-4f, 1.0f);
            float p = 1.0f / std::sqrt(cos2t);
            absorb = V(std::pow(tc.x, p), std::pow(tc.y, p), std::pow(tc.z, p));
        }
        f3 denom = max3(splat(1e-4f), 1.0f - (absorb * absorb) * (r1 * r1));
        c.thinT = (absorb * (t1 * t1)) / denom;
        c.thinR = r1 + c.thinT * (r1 * absorb);
        float thinRough =
            clampf(rough * std::sqrt(3.4f * (thinEta - 1.0f) * (thinEta - 0.5f) * (thinEta - 0.5f) /
                                     (thinEta * thinEta * thinEta)),
                   MIN_SPECULAR_ROUGHNESS, 1.0f);
        c.thinAlpha = thinRough * thinRough;
        c.thinScaleR = reflectionScale(t, rough, mu, splat(1.0f)).x;
        c.thinScaleT = reflectionScale(t, thinRough, mu, splat(1.0f)).x;
        float total = average(c.thinR) + average(c.thinT);
        c.thinTShare = total > 0.0f ? average(c.thinT) / total : 0.0f;
    }

    float schlickS = saturate((c.specE - f0) / std::fmax(1e-4f, 1.0f - f0));
    f3 metalE = lerp3(m.baseColor, splat(1.0f), schlickS);
    float glassTotal =
        thin ? average(c.thinR + c.thinT) : c.specE + (1.0f - c.specE) * average(c.glassTint);
    float dielectricE = lerp(c.specE, glassTotal, m.transmissionWeight);
    float baseAvg = average(c.baseWeight);
    c.wFuzz = exiting ? 0.0f : m.fuzzWeight * c.fuzzR * average(m.fuzzColor);
    c.wCoat = exiting ? 0.0f : c.fuzzAtten * m.coatWeight * c.coatE;
    c.wSpec = c.hasSpecular ? baseAvg * lerp(dielectricE, average(metalE), m.metalness) : 0.0f;
    c.wDiffuse = exiting ? 0.0f
                         : baseAvg * (1.0f - m.metalness) * (1.0f - m.transmissionWeight) *
                               (1.0f - c.specE) * average(m.baseColor);
    if (exiting) c.wSpec = 1.0f;
    c.wTotal = c.wFuzz + c.wCoat + c.wSpec + c.wDiffuse;

    f3 glassColor = thin ? c.thinR + c.thinT : splat(c.specE) + (1.0f - c.specE) * c.glassTint;
    if (exiting) {
        c.aovAlbedo = glassColor;
    } else {
        f3 dielectricColor = lerp3(splat(c.specE), glassColor, m.transmissionWeight);
        f3 specColor = c.hasSpecular ? lerp3(dielectricColor, metalE, m.metalness) : splat(0.0f);
        c.aovAlbedo = ((m.fuzzWeight * c.fuzzR) * m.fuzzColor +
                       splat(c.fuzzAtten * m.coatWeight * c.coatE)) +
                      c.baseWeight *
                          (specColor + (((1.0f - m.metalness) * (1.0f - m.transmissionWeight)) *
                                        (1.0f - c.specE)) *
                                           m.baseColor);
    }
    return c;
}

// openpbr.slang:596-608
float aovRecordFraction(const Closure& c, uint32_t lobe) {
    float alpha;
    if (lobe == LOBE_COAT) {
        alpha = c.coatAlpha;
    } else if (lobe == LOBE_SPECULAR) {
        alpha = c.alpha;
    } else if (lobe == LOBE_TRANSMISSION) {
        alpha = c.m.thinWalled != 0 ? c.thinAlpha : c.alpha;
    } else {
        return 1.0f;
    }
    return saturate(std::sqrt(alpha) / 0.15f);
}

// openpbr.slang:614-622
f3 aovPassTint(const Closure& c, uint32_t lobe) {
    if (lobe == LOBE_TRANSMISSION) return c.glassTint;
    if (lobe == LOBE_SPECULAR) return lerp3(splat(1.0f), c.m.baseColor, c.m.metalness);
    return splat(1.0f);
}

// openpbr.slang:629-636
float sheenDensity(const Closure& c, f3 wiStd) {
    float a = c.fuzzA, b = c.fuzzB;
    float lenSqr = (a * wiStd.x + b * wiStd.z) * (a * wiStd.x + b * wiStd.z) +
                   (a * wiStd.y) * (a * wiStd.y) + wiStd.z * wiStd.z;
    float ratio = a / lenSqr;
    return std::fmax(wiStd.z, 0.0f) * ratio * ratio / PI_OPENPBR;
}

// openpbr.slang:639-647
f3 alignToView(f3 v, f3 wo, bool inverse) {
    float len = std::sqrt(wo.x * wo.x + wo.y * wo.y);
    if (len < 1e-6f) return v;
    float cosPhi = wo.x / len;
    float sinPhi = (inverse ? -wo.y : wo.y) / len;
    return V(cosPhi * v.x + sinPhi * v.y, cosPhi * v.y - sinPhi * v.x, v.z);
}

// openpbr.slang:650-653
inline float reflectionCore(float alpha, f3 wi, f3 wo, f3 h) {
    return ggxD(alpha, h.z) * smithG1(alpha, wo.z) * smithG1(alpha, wi.z) / (4.0f * wo.z * wi.z);
}

// openpbr.slang:660-677
f3 baseReflection(const Closure& c, f3 wi, f3 wo, f3 h) {
    float core = reflectionCore(c.alpha, wi, wo, h);
    float cosVH = dot(h, wo);
    float fd = fresnelDielectric(c.eta, cosVH);
    if (c.exiting) return (splat(core * fd)) * c.glassScale;
    f3 conductor = fresnelSchlick(c.m.baseColor, cosVH) * c.conductorScale;
    f3 dielectric =
        lerp3(splat(fd * c.dielectricScale), fd * c.glassScale, c.m.transmissionWeight);
    if (c.m.thinWalled != 0) {
        dielectric = lerp3(splat(fd * c.dielectricScale), c.thinR * c.thinScaleR,
                           c.m.transmissionWeight);
    }
    return core * lerp3(dielectric, conductor, c.m.metalness);
}

// openpbr.slang:681-744
f3 evalBsdf(const Closure& c, f3 wi, f3 wo) {
    if (wo.z <= 0.0f || wi.z == 0.0f) return splat(0.0f);
    const Material& m = c.m;
    if (wi.z < 0.0f) {
        float glassShare = c.exiting ? 1.0f : (1.0f - m.metalness) * m.transmissionWeight;
        if (glassShare <= 0.0f || (c.exiting && m.thinWalled != 0)) return splat(0.0f);
        if (m.thinWalled != 0) {
            f3 flipped = V(wi.x, wi.y, -wi.z);
            f3 h = normalize(flipped + wo);
            float core = reflectionCore(c.thinAlpha, flipped, wo, h);
            return ((c.baseWeight * glassShare) * c.thinT) * c.thinScaleT * core;
        }
        f3 hU = wo + c.eta * wi;
        float len2 = dot(hU, hU);
        if (len2 < 1e-9f) return splat(0.0f);
        f3 h = hU / std::sqrt(len2);
        h = h.z >= 0.0f ? h : -h;
        float cosVH = dot(h, wo);
        float cosLH = dot(h, wi);
        if (cosVH <= 0.0f || cosLH >= 0.0f) return splat(0.0f);
        float f = fresnelDielectric(c.eta, cosVH);
        float core = ggxD(c.alpha, h.z) * smithG1(c.alpha, wo.z) * smithG1(c.alpha, wi.z) * cosVH *
                     (c.eta * c.eta) * std::fabs(cosLH) / (wo.z * std::fabs(wi.z) * len2);
        return ((((c.baseWeight * glassShare) * (1.0f - f)) * c.glassTint) * c.glassScale) * core;
    }

    f3 result = splat(0.0f);
    if (c.fuzzR > 0.0f) {
        float density = sheenDensity(c, alignToView(wi, wo, false));
        result = result + ((m.fuzzWeight * m.fuzzColor) * c.fuzzR) * density / wi.z;
    }
    f3 h = normalize(wi + wo);
    if (m.coatWeight > 0.0f && !c.exiting) {
        float coatF = fresnelDielectric(m.coatIor, dot(h, wo));
        result = result + splat(c.fuzzAtten * m.coatWeight * coatF * c.coatScale *
                                reflectionCore(c.coatAlpha, wi, wo, h));
    }
    f3 base = splat(0.0f);
    if (c.hasSpecular) base = baseReflection(c, wi, wo, h);
    if (!c.exiting) {
        base = base + (((1.0f - m.metalness) * (1.0f - m.transmissionWeight) * (1.0f - c.specE)) *
                       eonBrdf(m.baseColor, m.baseRoughness, wi, wo));
    }
    return result + c.baseWeight * base;
}

// openpbr.slang:748-752
inline float reflectProbability(const Closure& c, float fresnel) {
    float glassShare = c.exiting ? 1.0f : (1.0f - c.m.metalness) * c.m.transmissionWeight;
    float transmit = c.m.thinWalled != 0 ? c.thinTShare : 1.0f - fresnel;
    return 1.0f - glassShare * transmit;
}

// openpbr.slang:756-808
float bsdfPdf(const Closure& c, f3 wi, f3 wo) {
    if (wo.z <= 0.0f || wi.z == 0.0f || c.wTotal <= 0.0f) return 0.0f;
    if (wi.z < 0.0f) {
        if (c.wSpec <= 0.0f) return 0.0f;
        float pdf = 0.0f;
        if (c.m.thinWalled != 0) {
            f3 flipped = V(wi.x, wi.y, -wi.z);
            f3 h = normalize(flipped + wo);
            pdf = vndfReflectPdf(c.thinAlpha, wo, h) * (1.0f - reflectProbability(c, 0.0f));
        } else {
            f3 hU = wo + c.eta * wi;
            float len2 = dot(hU, hU);
            if (len2 < 1e-9f) return 0.0f;
            f3 h = hU / std::sqrt(len2);
            h = h.z >= 0.0f ? h : -h;
            float cosVH = dot(h, wo);
            float cosLH = dot(h, wi);
            if (cosVH <= 0.0f || cosLH >= 0.0f) return 0.0f;
            float f = fresnelDielectric(c.eta, cosVH);
            float vndf = smithG1(c.alpha, wo.z) * ggxD(c.alpha, h.z) * cosVH / wo.z;
            float jacobian = (c.eta * c.eta) * std::fabs(cosLH) / len2;
            pdf = vndf * jacobian * (1.0f - reflectProbability(c, f));
        }
        return c.wSpec * pdf / c.wTotal;
    }
    f3 h = normalize(wi + wo);
    float pdf = 0.0f;
    if (c.wFuzz > 0.0f) pdf += c.wFuzz * sheenDensity(c, alignToView(wi, wo, false));
    if (c.wCoat > 0.0f) pdf += c.wCoat * vndfReflectPdf(c.coatAlpha, wo, h);
    if (c.wSpec > 0.0f) {
        float f = fresnelDielectric(c.eta, dot(h, wo));
        pdf += c.wSpec * vndfReflectPdf(c.alpha, wo, h) * reflectProbability(c, f);
    }
    if (c.wDiffuse > 0.0f) pdf += c.wDiffuse * wi.z / PI_OPENPBR;
    return pdf / c.wTotal;
}

struct BsdfSample {
    f3 wi;
    float pdf;
    uint32_t lobe;
};

// openpbr.slang:821-884
BsdfSample sampleBsdf(const Closure& c, f3 wo, float uLobe, f2 u) {
    BsdfSample result{V(0.0f, 0.0f, 1.0f), 0.0f, LOBE_DIFFUSE};
    if (c.wTotal <= 0.0f) return result;
    float pick = uLobe * c.wTotal;
    f3 wi = splat(0.0f);
    if (pick < c.wFuzz) {
        result.lobe = LOBE_FUZZ;
        float r = std::sqrt(u.x);
        float phi = 2.0f * PI_OPENPBR * u.y;
        f2 disk{r * std::cos(phi), r * std::sin(phi)};
        float diskZ = std::sqrt(std::fmax(0.0f, 1.0f - (disk.x * disk.x + disk.y * disk.y)));
        f3 wiStd = normalize(V(disk.x - diskZ * c.fuzzB, disk.y, diskZ * c.fuzzA));
        wi = alignToView(wiStd, wo, true);
    } else if (pick < c.wFuzz + c.wCoat) {
        result.lobe = LOBE_COAT;
        f3 h = sampleVndf(wo, c.coatAlpha, u);
        wi = (2.0f * dot(wo, h)) * h - wo;
    } else if (pick < c.wFuzz + c.wCoat + c.wSpec) {
        float uBranch = (pick - c.wFuzz - c.wCoat) / std::fmax(1e-6f, c.wSpec);
        if (c.m.thinWalled != 0) {
            if (uBranch < reflectProbability(c, 0.0f)) {
                result.lobe = LOBE_SPECULAR;
                f3 h = sampleVndf(wo, c.alpha, u);
                wi = (2.0f * dot(wo, h)) * h - wo;
            } else {
                result.lobe = LOBE_TRANSMISSION;
                f3 h = sampleVndf(wo, c.thinAlpha, u);
                f3 reflected = (2.0f * dot(wo, h)) * h - wo;
                wi = V(reflected.x, reflected.y, -reflected.z);
            }
        } else {
            f3 h = sampleVndf(wo, c.alpha, u);
            float f = fresnelDielectric(c.eta, dot(h, wo));
            if (uBranch < reflectProbability(c, f)) {
                result.lobe = LOBE_SPECULAR;
                wi = (2.0f * dot(wo, h)) * h - wo;
            } else {
                result.lobe = LOBE_TRANSMISSION;
                f4 refracted = refractDirection(wo, h, c.eta);
                if (refracted.w == 0.0f) return result;
                wi = V(refracted.x, refracted.y, refracted.z);
            }
        }
    } else {
        result.lobe = LOBE_DIFFUSE;
        f4 s = sampleCosineHemisphere(u);
        wi = V(s.x, s.y, s.z);
    }
    result.wi = wi;
    result.pdf = bsdfPdf(c, wi, wo);
    return result;
}

// openpbr.slang:888-909 (Duff et al. 2017)
struct Frame {
    f3 t, b, n;
    explicit Frame(f3 nn) : n(nn) {
        float s = nn.z >= 0.0f ? 1.0f : -1.0f;
        float a = -1.0f / (s + nn.z);
        float m = nn.x * nn.y * a;
        t = V(1.0f + s * nn.x * nn.x * a, s * m, -s * nn.x);
        b = V(m, s + nn.y * nn.y * a, -nn.y);
    }
    f3 toLocal(f3 v) const { return V(dot(v, t), dot(v, b), dot(v, n)); }
    f3 toWorld(f3 v) const { return (v.x * t + v.y * b) + v.z * n; }
};

// ------------------------------------------------------------ scene

constexpr uint32_t MASK_CAMERA = 0x01;  // scene/mod.rs:54-57
constexpr uint32_t MASK_ALL = 0xFF;
constexpr uint32_t MEDIUM_NONE = 0xffffff;  // pathstate.slang:188

struct Instance {
    float o2w[3][4];
    float w2o[3][4];
    uint32_t light;  // first light record or LIGHT_NONE
    uint32_t mesh;
    uint32_t mask;
    bool opaque;
    int32_t log2_transmittance;  // log2_transmittance_fixed(opacity) for a non-opaque instance
};

struct MeshData {
    std::vector<float> positions, normals, uvs;
    std::vector<uint32_t> indices;
};

// One flattened world-space triangle for traversal.
struct WorldTri {
    f3 v0, v1, v2;
    uint32_t instance, primitive;
};

struct Node {  // binary BVH node
    float lo[3], hi[3];
    uint32_t left;   // child index, or first triangle for a leaf
    uint32_t count;  // 0 for internal nodes
};

// glam Mat4::transform_point3 over the rows of a 3x4 transform: the
// columns accumulate x, then y, then z, and the translation adds last.
inline f3 transform_point(const float r[3][4], f3 p) {
    return V(((r[0][0] * p.x + r[0][1] * p.y) + r[0][2] * p.z) + r[0][3],
             ((r[1][0] * p.x + r[1][1] * p.y) + r[1][2] * p.z) + r[1][3],
             ((r[2][0] * p.x + r[2][1] * p.y) + r[2][2] * p.z) + r[2][3]);
}

// The inverse of an affine 3x4 transform, in f64, rounded to f32.
void invert_affine(const float m[3][4], float out[3][4]) {
    double a[3][3], t[3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) a[i][j] = m[i][j];
        t[i] = m[i][3];
    }
    double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
    double inv[3][3];
    inv[0][0] = c00 / det;
    inv[1][0] = c01 / det;
    inv[2][0] = c02 / det;
    inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) out[i][j] = float(inv[i][j]);
        out[i][3] = float(-(inv[i][0] * t[0] + inv[i][1] * t[1] + inv[i][2] * t[2]));
    }
}

}  // namespace

struct orc_scene {
    std::vector<MeshData> meshes;
    std::vector<Instance> instances;
    std::vector<Material> materials;
    std::vector<LightRec> lights;
    std::vector<WorldTri> tris;
    std::vector<Node> nodes;
    EnvTables env;
    uint32_t env_w = 0, env_h = 0;
    std::vector<float> env_texels;
    float select_prob = 0.0f;
    std::vector<float> tables;
    cn_camera camera;
};

namespace {

// ------------------------------------------------------------- BVH

void node_bounds(const orc_scene& s, const uint32_t* ids, uint32_t n, Node& node) {
    for (int k = 0; k < 3; ++k) {
        node.lo[k] = INFINITY;
        node.hi[k] = -INFINITY;
    }
    for (uint32_t i = 0; i < n; ++i) {
        const WorldTri& t = s.tris[ids[i]];
        const f3 v[3] = {t.v0, t.v1, t.v2};
        for (const f3& p : v) {
            const float c[3] = {p.x, p.y, p.z};
            for (int k = 0; k < 3; ++k) {
                node.lo[k] = std::fmin(node.lo[k], c[k]);
                node.hi[k] = std::fmax(node.hi[k], c[k]);
            }
        }
    }
    // Grow by a relative margin so the f32 slab test stays conservative.
    for (int k = 0; k < 3; ++k) {
        float pad = 1e-5f * std::fmax(std::fabs(node.lo[k]), std::fabs(node.hi[k])) + 1e-30f;
        node.lo[k] -= pad;
        node.hi[k] += pad;
    }
}

// A plain median-split binary BVH: the oracle's own acceleration, whose
// only contract is to return the same closest hit a brute-force loop would.
void build_bvh(orc_scene& s) {
    std::vector<uint32_t> ids(s.tris.size());
    for (uint32_t i = 0; i < ids.size(); ++i) ids[i] = i;
    std::vector<f3> centroid(s.tris.size());
    for (size_t i = 0; i < s.tris.size(); ++i) {
        centroid[i] = (s.tris[i].v0 + s.tris[i].v1 + s.tris[i].v2) * (1.0f / 3.0f);
    }
    struct Task {
        uint32_t node, begin, end;
    };
    s.nodes.clear();
    s.nodes.push_back(Node{});
    std::vector<Task> stack{{0, 0, uint32_t(ids.size())}};
    while (!stack.empty()) {
        Task task = stack.back();
        stack.pop_back();
        Node node{};
        node_bounds(s, ids.data() + task.begin, task.end - task.begin, node);
        uint32_t n = task.end - task.begin;
        if (n <= 4) {
            node.left = task.begin;
            node.count = n;
            s.nodes[task.node] = node;
            continue;
        }
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = task.begin; i < task.end; ++i) {
            const f3& c = centroid[ids[i]];
            const float cc[3] = {c.x, c.y, c.z};
            for (int k = 0; k < 3; ++k) {
                clo[k] = std::fmin(clo[k], cc[k]);
                chi[k] = std::fmax(chi[k], cc[k]);
            }
        }
        int axis = 0;
        for (int k = 1; k < 3; ++k) {
            if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
        }
        uint32_t mid = task.begin + n / 2;
        std::nth_element(ids.begin() + task.begin, ids.begin() + mid, ids.begin() + task.end,
                         [&](uint32_t a, uint32_t b) {
                             const float ca[3] = {centroid[a].x, centroid[a].y, centroid[a].z};
                             const float cb[3] = {centroid[b].x, centroid[b].y, centroid[b].z};
                             return ca[axis] < cb[axis] || (ca[axis] == cb[axis] && a < b);
                         });
        uint32_t left = uint32_t(s.nodes.size());
        s.nodes.push_back(Node{});
        s.nodes.push_back(Node{});
        node.left = left;
        node.count = 0;
        s.nodes[task.node] = node;
        stack.push_back({left, task.begin, mid});
        stack.push_back({left + 1, mid, task.end});
    }
    // Reorder triangles into leaf order.
    std::vector<WorldTri> ordered(s.tris.size());
    for (size_t i = 0; i < ids.size(); ++i) ordered[i] = s.tris[ids[i]];
    s.tris.swap(ordered);
}

struct Ray {
    f3 o, d;
    float tmax;
};

// Watertight ray/triangle test (Woop, Benthin & Wald, JCGT 2(1), 2013):
// on a hit, t and the barycentrics (b1, b2) of v1 and v2.
struct WatertightRay {
    int kx, ky, kz;
    float Sx, Sy, Sz;
    f3 o;
};

WatertightRay prepare(const Ray& r) {
    WatertightRay w;
    float ad[3] = {std::fabs(r.d.x), std::fabs(r.d.y), std::fabs(r.d.z)};
    int kz = 0;
    if (ad[1] > ad[kz]) kz = 1;
    if (ad[2] > ad[kz]) kz = 2;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    const float d[3] = {r.d.x, r.d.y, r.d.z};
    if (d[kz] < 0.0f) std::swap(kx, ky);
    w.kx = kx;
    w.ky = ky;
    w.kz = kz;
    w.Sx = d[kx] / d[kz];
    w.Sy = d[ky] / d[kz];
    w.Sz = 1.0f / d[kz];
    w.o = r.o;
    return w;
}

inline float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

bool intersect_triangle(const WatertightRay& w, f3 v0, f3 v1, f3 v2, float tmax, float* t_out,
                        float* b1, float* b2) {
    f3 A = v0 - w.o, B = v1 - w.o, C = v2 - w.o;
    float Az = comp(A, w.kz), Bz = comp(B, w.kz), Cz = comp(C, w.kz);
    float Ax = comp(A, w.kx) - w.Sx * Az, Ay = comp(A, w.ky) - w.Sy * Az;
    float Bx = comp(B, w.kx) - w.Sx * Bz, By = comp(B, w.ky) - w.Sy * Bz;
    float Cx = comp(C, w.kx) - w.Sx * Cz, Cy = comp(C, w.ky) - w.Sy * Cz;
    float U = Cx * By - Cy * Bx;
    float Vv = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || Vv == 0.0f || W == 0.0f) {
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        Vv = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.0f || Vv < 0.0f || W < 0.0f) && (U > 0.0f || Vv > 0.0f || W > 0.0f)) return false;
    float det = U + Vv + W;
    if (det == 0.0f) return false;
    Az = w.Sz * Az;
    Bz = w.Sz * Bz;
    Cz = w.Sz * Cz;
    float T = U * Az + Vv * Bz + W * Cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= 0.0f) || !(t <= tmax)) return false;
    *t_out = t;
    *b1 = Vv * rcp;
    *b2 = W * rcp;
    return true;
}

struct RayCtx {
    f3 inv;
    WatertightRay w;
};

inline bool box_hit(const Node& n, const Ray& r, const f3& inv, float tmax, float* tnear) {
    float t0 = 0.0f, t1 = tmax;
    const float o[3] = {r.o.x, r.o.y, r.o.z};
    const float iv[3] = {inv.x, inv.y, inv.z};
    for (int k = 0; k < 3; ++k) {
        float a = (n.lo[k] - o[k]) * iv[k];
        float b = (n.hi[k] - o[k]) * iv[k];
        float lo = std::fmin(a, b), hi = std::fmax(a, b);
        t0 = std::fmax(t0, lo);
        t1 = std::fmin(t1, hi * 1.0000004f);
    }
    *tnear = t0;
    return t0 <= t1;
}

inline f3 safe_inverse(f3 d) {
    auto inv = [](float x) {
        float ax = std::fabs(x);
        float s = ax < 1e-20f ? std::copysign(1e-20f, x) : x;
        return 1.0f / s;
    };
    return V(inv(d.x), inv(d.y), inv(d.z));
}

struct HitRecord {
    bool hit;
    float t;
    uint32_t instance, primitive;
    float b1, b2;
};

// Closest hit within [0, ray.tmax] among triangles whose instance mask
// meets `mask`. `candidate` decides each crossing of a non-opaque instance
// (shaders/intersect.slang:172-188): return true to commit it.
template <typename Candidate>
HitRecord closest_hit(const orc_scene& s, const Ray& ray, uint32_t mask, Candidate&& candidate) {
    HitRecord best{false, ray.tmax, 0, 0, 0.0f, 0.0f};
    uint64_t best_key = ~0ull;
    f3 inv = safe_inverse(ray.d);
    WatertightRay w = prepare(ray);
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const Node& n = s.nodes[stack[--sp]];
        float tn;
        if (!box_hit(n, ray, inv, best.t, &tn)) continue;
        if (n.count > 0) {
            for (uint32_t i = n.left; i < n.left + n.count; ++i) {
                const WorldTri& tri = s.tris[i];
                const Instance& inst = s.instances[tri.instance];
                if ((inst.mask & mask) == 0) continue;
                float t, b1, b2;
                if (!intersect_triangle(w, tri.v0, tri.v1, tri.v2, best.t, &t, &b1, &b2)) continue;
                uint64_t key = (uint64_t(tri.instance) << 32) | tri.primitive;
                if (best.hit && t == best.t && key > best_key) continue;
                if (!inst.opaque && !candidate(tri.instance, tri.primitive, t)) continue;
                best = {true, t, tri.instance, tri.primitive, b1, b2};
                best_key = key;
            }
        } else {
            stack[sp++] = n.left + 1;
            stack[sp++] = n.left;
        }
    }
    return best;
}

// -------------------------------------------------------- environment

struct EnvSample {
    f3 direction, radiance;
    float pdf;
};

// Exact fp32 bilinear, REPEAT in u and CLAMP in v (gpu/image.rs:95-103,
// :263-270: the sampler the reference binds), texel centers at +0.5.
f3 env_lookup(const orc_scene& s, float u, float v) {
    const int w = int(s.env_w), h = int(s.env_h);
    float x = u * float(w) - 0.5f;
    float y = v * float(h) - 0.5f;
    float fx = std::floor(x), fy = std::floor(y);
    float ax = x - fx, ay = y - fy;
    int x0 = int(fx), y0 = int(fy);
    int x1 = x0 + 1, y1 = y0 + 1;
    x0 = ((x0 % w) + w) % w;
    x1 = ((x1 % w) + w) % w;
    y0 = std::min(std::max(y0, 0), h - 1);
    y1 = std::min(std::max(y1, 0), h - 1);
    auto texel = [&](int xx, int yy) {
        const float* p = &s.env_texels[4 * (size_t(yy) * w + xx)];
        return V(p[0], p[1], p[2]);
    };
    f3 top = texel(x0, y0) * (1.0f - ax) + texel(x1, y0) * ax;
    f3 bottom = texel(x0, y1) * (1.0f - ax) + texel(x1, y1) * ax;
    return top * (1.0f - ay) + bottom * ay;
}

// environment.slang:74-78
inline f2 uvFromDirection(f3 d) {
    float u = 0.5f + std::atan2(d.x, -d.z) / (2.0f * PI_SLANG);
    float v = std::acos(clampf(d.y, -1.0f, 1.0f)) / PI_SLANG;
    return {u, v};
}

// environment.slang:80-85
inline f3 directionFromUv(f2 uv) {
    float phi = 2.0f * PI_SLANG * (uv.x - 0.5f);
    float theta = PI_SLANG * uv.y;
    float sinTheta = std::sin(theta);
    return V(sinTheta * std::sin(phi), std::cos(theta), -sinTheta * std::cos(phi));
}

// environment.slang:88-90
inline f3 env_radiance(const orc_scene& s, f3 d) {
    f2 uv = uvFromDirection(d);
    return env_lookup(s, uv.x, uv.y);
}

// environment.slang:41-58
struct CdfSlot {
    uint32_t index;
    float position;
};
CdfSlot sampleCdf(const float* cdf, uint32_t count, float u) {
    uint32_t lo = 0, hi = count;
    while (lo + 1 < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid;
    }
    float width = cdf[lo + 1] - cdf[lo];
    float frac = width > 0.0f ? (u - cdf[lo]) / width : 0.0f;
    return {lo, (float(lo) + frac) / float(count)};
}

// environment.slang:94-106
EnvSample env_sample(const orc_scene& s, f2 u) {
    CdfSlot row = sampleCdf(s.env.marginal.data(), s.env_h, u.x);
    CdfSlot col = sampleCdf(s.env.conditional.data() + row.index * (s.env_w + 1), s.env_w, u.y);
    f2 uv{col.position, row.position};
    float sinTheta = std::sin(PI_SLANG * uv.y);
    EnvSample out;
    out.direction = directionFromUv(uv);
    out.radiance = env_lookup(s, uv.x, uv.y);
    out.pdf = sinTheta > 0.0f ? s.env.pdfs[row.index * s.env_w + col.index] /
                                    (2.0f * PI_SLANG * PI_SLANG * sinTheta)
                              : 0.0f;
    return out;
}

// environment.slang:110-119
float env_pdf(const orc_scene& s, f3 d) {
    float sinTheta = std::sqrt(std::fmax(0.0f, 1.0f - d.y * d.y));
    if (sinTheta <= 0.0f) return 0.0f;
    f2 uv = uvFromDirection(d);
    uint32_t col = std::min(uint32_t(uv.x * float(s.env_w)), s.env_w - 1);
    uint32_t row = std::min(uint32_t(uv.y * float(s.env_h)), s.env_h - 1);
    return s.env.pdfs[row * s.env_w + col] / (2.0f * PI_SLANG * PI_SLANG * sinTheta);
}

// --------------------------------------------------------- lights.slang

inline float powerHeuristic(float p, float q) { return (p * p) / (p * p + q * q); }  // :42-44

struct LightConnection {
    f3 direction;
    float tmax;
    f3 radiance;
    float pdf;
    bool delta;
    uint32_t instance, primitive;
};

// lights.slang:93-160
LightConnection sampleLights(const orc_scene& s, float u, f2 uPoint, f3 from) {
    uint32_t count = uint32_t(s.lights.size());
    float scaled = u * float(count);
    uint32_t slot = std::min(uint32_t(scaled), count - 1);
    const LightRec* light = &s.lights[slot];
    if (scaled - float(slot) >= light->alias_threshold) light = &s.lights[light->alias_index];

    LightConnection c;
    c.pdf = 0.0f;
    c.direction = splat(0.0f);
    c.tmax = 1e30f;
    c.radiance = light->emission;
    c.delta = light->kind != KIND_TRIANGLE;
    c.instance = light->instance;
    c.primitive = light->primitive;
    switch (light->kind) {
        case KIND_TRIANGLE: {
            float su = std::sqrt(uPoint.x);
            f2 bary{1.0f - su, uPoint.y * su};
            f3 position = (light->anchor + bary.x * light->edge1) + bary.y * light->edge2;
            f3 delta = position - from;
            float r2 = dot(delta, delta);
            f3 normal = normalize(cross(light->edge1, light->edge2));
            f3 direction = delta * rsqrt(r2);
            float cosLight = -dot(normal, direction);
            if (r2 > 0.0f && cosLight > 0.0f && light->pdf > 0.0f) {
                c.direction = direction;
                c.pdf = light->pdf * r2 / cosLight;
                c.tmax = std::sqrt(r2) * 1.001f;
            }
            break;
        }
        case KIND_DISTANT:
            c.direction = -light->anchor;
            c.pdf = light->pdf;
            break;
        default: {
            f3 delta = light->anchor - from;
            float r2 = dot(delta, delta);
            if (r2 > 0.0f && light->pdf > 0.0f) {
                float r = std::sqrt(r2);
                c.direction = delta / r;
                c.radiance = light->emission / r2;
                c.pdf = light->pdf;
                c.tmax = r;
            }
            break;
        }
    }
    return c;
}

// ------------------------------------------------- shade_surface.slang

struct SpawnPoint {
    f3 position, below, normal, shading;
    bool frontFace;
};

inline f3 fetch3(const std::vector<float>& a, uint32_t v) {
    return V(a[3 * v], a[3 * v + 1], a[3 * v + 2]);
}

// shade_surface.slang:96-181 (van Antwerpen 2023 offsets).
SpawnPoint spawnPoint(const orc_scene& s, uint32_t instance, uint32_t primitive, f2 bary,
                      f3 rayDirection) {
    const Instance& inst = s.instances[instance];
    const MeshData& mesh = s.meshes[inst.mesh];
    uint32_t base = 3 * primitive;
    uint32_t i0 = mesh.indices[base], i1 = mesh.indices[base + 1], i2 = mesh.indices[base + 2];
    f3 v0 = fetch3(mesh.positions, i0), v1 = fetch3(mesh.positions, i1),
       v2 = fetch3(mesh.positions, i2);
    const float(*o2w)[4] = inst.o2w;
    const float(*w2o)[4] = inst.w2o;

    f3 edge1 = v1 - v0;
    f3 edge2 = v2 - v0;
    f3 objPosition = v0 + fma3(splat(bary.x), edge1, bary.y * edge2);
    f3 objNormal = cross(edge1, edge2);

    f3 wldPosition = V(
        o2w[0][3] + std::fmaf(o2w[0][0], objPosition.x,
                              std::fmaf(o2w[0][1], objPosition.y, o2w[0][2] * objPosition.z)),
        o2w[1][3] + std::fmaf(o2w[1][0], objPosition.x,
                              std::fmaf(o2w[1][1], objPosition.y, o2w[1][2] * objPosition.z)),
        o2w[2][3] + std::fmaf(o2w[2][0], objPosition.x,
                              std::fmaf(o2w[2][1], objPosition.y, o2w[2][2] * objPosition.z)));
    f3 r0 = V(w2o[0][0], w2o[0][1], w2o[0][2]);
    f3 r1 = V(w2o[1][0], w2o[1][1], w2o[1][2]);
    f3 r2 = V(w2o[2][0], w2o[2][1], w2o[2][2]);
    f3 wldNormal = (r0 * objNormal.x + r1 * objNormal.y) + r2 * objNormal.z;
    float wldScale = rsqrt(dot(wldNormal, wldNormal));
    wldNormal = wldNormal * wldScale;

    const float c0 = 5.9604644775390625e-8f;
    const float c1 = 1.788139769587360206060111522674560546875e-7f;
    const float c2 = 1.19209317972490680404007434844970703125e-7f;

    f3 extent3 = (abs3(edge1) + abs3(edge2)) + abs3(edge1 - edge2);
    float extent = std::fmax(std::fmax(extent3.x, extent3.y), extent3.z);
    f3 objErr = fma3(splat(c0), abs3(v0), splat(c1 * extent));

    f3 q0 = V(o2w[0][0], o2w[0][1], o2w[0][2]);
    f3 q1 = V(o2w[1][0], o2w[1][1], o2w[1][2]);
    f3 q2 = V(o2w[2][0], o2w[2][1], o2w[2][2]);
    f3 aobj = abs3(objPosition);
    f3 wldErr = fma3(splat(c1), V(dot(abs3(q0), aobj), dot(abs3(q1), aobj), dot(abs3(q2), aobj)),
                     c2 * abs3(V(o2w[0][3], o2w[1][3], o2w[2][3])));
    f3 awld = abs3(wldPosition);
    objErr = fma3(splat(c2),
                  V(dot(abs3(r0), awld) + std::fabs(w2o[0][3]),
                    dot(abs3(r1), awld) + std::fabs(w2o[1][3]),
                    dot(abs3(r2), awld) + std::fabs(w2o[2][3])),
                  objErr);

    float wldOffset = dot(wldErr, abs3(wldNormal));
    float objOffset = dot(objErr, abs3(objNormal));
    wldOffset = std::fmaf(wldScale, objOffset, wldOffset);

    bool frontFace = dot(wldNormal, rayDirection) < 0.0f;
    if (dot(wldNormal, rayDirection) > 0.0f) wldNormal = -wldNormal;

    f3 n0 = fetch3(mesh.normals, i0), n1 = fetch3(mesh.normals, i1), n2 = fetch3(mesh.normals, i2);
    f3 objShading = (n0 + bary.x * (n1 - n0)) + bary.y * (n2 - n0);
    f3 shading = normalize((r0 * objShading.x + r1 * objShading.y) + r2 * objShading.z);
    if (dot(shading, wldNormal) < 0.0f) shading = -shading;

    SpawnPoint sp;
    sp.position = fma3(splat(wldOffset), wldNormal, wldPosition);
    sp.below = fma3(splat(-wldOffset), wldNormal, wldPosition);
    sp.normal = wldNormal;
    sp.shading = shading;
    sp.frontFace = frontFace;
    return sp;
}

// The per-pixel accumulators one wave writes (wavefront.rs:733-782).
struct PixelOut {
    f4 radiance{0, 0, 0, 0};
    f4 albedo{0, 0, 0, 0};
    f4 normal{0, 0, 0, 0};
    float depth = 0.0f;
};

inline void add4(f4& a, f3 v, float w) {
    a.x += v.x;
    a.y += v.y;
    a.z += v.z;
    a.w += w;
}

constexpr uint32_t ROULETTE_START = 3;  // shade_surface.slang:53

// One path of one wave: raygen.slang:64-112, then per bounce
// intersect.slang:45-88, shade_miss.slang:27-75 or
// shade_surface.slang:233-521, then trace_shadow.slang:27-72 — the same
// per-pixel order of adds the stage barriers give the reference.
void trace_path(const orc_scene& s, const Basis& basis, float aperture, uint32_t width,
                uint32_t height, uint32_t pixel, uint32_t sample, uint32_t maxBounces,
                uint32_t lightSampling, bool aovOn, PixelOut& out) {
    // ---- raygen (raygen.slang:76-110)
    f3 camPos = V(s.camera.position[0], s.camera.position[1], s.camera.position[2]);
    f2 jitter = sample_2d(pixel, sample, CAMERA_JITTER);
    float xy_x = float(pixel % width) + jitter.x;
    float xy_y = float(pixel / width) + jitter.y;
    float uvx = xy_x / float(width), uvy = xy_y / float(height);
    float ndcx = 2.0f * uvx - 1.0f, ndcy = 1.0f - 2.0f * uvy;
    f3 target = (basis.forward + ndcx * basis.right) + ndcy * basis.up;
    f3 origin = camPos;
    f3 direction = normalize(target);
    if (aperture > 0.0f) {
        // concentricDisk, raygen.slang:47-62
        f2 u = sample_2d(pixel, sample, CAMERA_LENS);
        float ox = 2.0f * u.x - 1.0f, oy = 2.0f * u.y - 1.0f;
        f2 disk{0.0f, 0.0f};
        if (!(ox == 0.0f && oy == 0.0f)) {
            float r, theta;
            if (std::fabs(ox) > std::fabs(oy)) {
                r = ox;
                theta = (PI_SLANG / 4.0f) * (oy / ox);
            } else {
                r = oy;
                theta = (PI_SLANG / 2.0f) - (PI_SLANG / 4.0f) * (ox / oy);
            }
            disk = {r * std::cos(theta), r * std::sin(theta)};
        }
        disk = {aperture * disk.x, aperture * disk.y};
        f3 offset = disk.x * normalize(basis.right) + disk.y * normalize(basis.up);
        origin = origin + offset;
        direction = normalize(target - offset);
    }
    float cosForward = dot(direction, normalize(basis.forward));
    f4 throughput{1.0f, 1.0f, 1.0f, 0.0f};
    uint32_t state = 0;
    f3 guideState = splat(1.0f);

    for (uint32_t bounce = 0; bounce < maxBounces; ++bounce) {
        // ---- intersect
        Ray ray{origin, direction, 1e30f};
        uint32_t mask = bounce == 0 ? MASK_CAMERA : MASK_ALL;
        HitRecord hit = closest_hit(s, ray, mask, [&](uint32_t inst, uint32_t prim, float) {
            return crossing_uniform(pixel, sample, bounce, inst, prim) < s.materials[inst].opacity;
        });

        if (!hit.hit) {
            // ---- shade_miss
            float weight = 1.0f;
            if (throughput.w > 0.0f) {
                if (lightSampling == CN_LIGHT_SAMPLING_NEE) {
                    weight = 0.0f;
                } else if (lightSampling == CN_LIGHT_SAMPLING_MIS) {
                    float neePdf = s.select_prob * env_pdf(s, direction);
                    weight = powerHeuristic(throughput.w, neePdf);
                }
            }
            if (aovOn) {
                bool cameraRay = throughput.w == 0.0f;
                if (cameraRay) out.depth += INFINITY;
                f3 g = cameraRay ? splat(1.0f) : guideState;
                add4(out.albedo, g, 0.0f);
            }
            f3 contribution = (V(throughput.x, throughput.y, throughput.z) *
                               env_radiance(s, direction)) *
                              weight;
            add4(out.radiance, contribution, 1.0f);
            return;
        }

        // ---- shade_surface
        const Instance& inst = s.instances[hit.instance];
        const Material& material = s.materials[hit.instance];
        SpawnPoint spawn = spawnPoint(s, hit.instance, hit.primitive, {hit.b1, hit.b2}, direction);

        uint32_t medium = bounce == 0 ? MEDIUM_NONE : (state & 0xffffff);
        if (medium != MEDIUM_NONE) {
            const Material& interior = s.materials[medium];
            if (interior.transmissionDepth > 0.0f) {
                f3 sigma = (-log3(clamp3(interior.transmissionColor, 1e-4f, 1.0f))) /
                           interior.transmissionDepth;
                f3 att = exp3(-sigma * length(origin - spawn.position));
                throughput.x *= att.x;
                throughput.y *= att.y;
                throughput.z *= att.z;
            }
        }
        f3 T = V(throughput.x, throughput.y, throughput.z);

        f3 contribution = splat(0.0f);
        if (inst.light != LIGHT_NONE && spawn.frontFace) {
            float weight = 1.0f;
            if (throughput.w > 0.0f) {
                if (lightSampling == CN_LIGHT_SAMPLING_NEE) {
                    weight = 0.0f;
                } else if (lightSampling == CN_LIGHT_SAMPLING_MIS) {
                    f3 delta = spawn.position - origin;
                    float cosLight = std::fabs(dot(spawn.normal, direction));
                    float neePdf = (1.0f - s.select_prob) * s.lights[inst.light + hit.primitive].pdf *
                                   dot(delta, delta) / cosLight;
                    weight = powerHeuristic(throughput.w, neePdf);
                }
            }
            contribution = (T * material.emission) * weight;
        }

        f3 shading = spawn.shading;
        Frame frame(shading);
        f3 wo = frame.toLocal(-direction);
        Closure closure = resolveClosure(s.tables.data(), material, wo, medium == hit.instance);

        f3 guide = splat(0.0f);
        if (aovOn) {
            guide = bounce == 0 ? splat(1.0f) : guideState;
            if (bounce == 0) out.depth += length(origin - spawn.position) * cosForward;
        }
        f3 aovAlbedo = clamp3(closure.aovAlbedo, 0.0f, 1.0f);
        auto recordGuides = [&](f3 w) {
            add4(out.albedo, w * aovAlbedo, 0.0f);
            add4(out.normal, w * shading, 0.0f);
        };

        if (bounce + 1 >= maxBounces) {
            if (aovOn && any_pos(guide)) recordGuides(guide);
            add4(out.radiance, contribution, 1.0f);
            return;
        }

        // Next-event estimation (shade_surface.slang:363-446).
        bool haveShadow = false;
        f3 shadowOrigin{}, shadowDir{}, shadowContribution{};
        float shadowTmax = 0.0f;
        uint32_t shadowLight = LIGHT_NONE, shadowPrim = 0;
        float envProb = s.select_prob;
        uint32_t lightCount = uint32_t(s.lights.size());
        if (lightSampling != CN_LIGHT_SAMPLING_BSDF && (lightCount > 0 || envProb > 0.0f)) {
            float uLight = sample_1d(pixel, sample, bounce_dimension(bounce, NEE_LIGHT));
            f2 uPoint = sample_2d(pixel, sample, bounce_dimension(bounce, NEE_POINT));
            f3 unshadowed = splat(0.0f);
            if (uLight < envProb) {
                EnvSample cand = env_sample(s, uPoint);
                f3 wiLight = frame.toLocal(cand.direction);
                bool throughSurface = dot(cand.direction, spawn.normal) < 0.0f;
                if (cand.pdf > 0.0f && wiLight.z != 0.0f && (wiLight.z < 0.0f) == throughSurface) {
                    float neePdf = envProb * cand.pdf;
                    float scatterPdf = bsdfPdf(closure, wiLight, wo);
                    float weight = lightSampling == CN_LIGHT_SAMPLING_MIS
                                       ? powerHeuristic(neePdf, scatterPdf)
                                       : 1.0f;
                    unshadowed = (((T * cand.radiance) * evalBsdf(closure, wiLight, wo)) *
                                  std::fabs(wiLight.z)) *
                                 (weight / neePdf);
                    shadowOrigin = throughSurface ? spawn.below : spawn.position;
                    shadowLight = LIGHT_NONE;
                    shadowPrim = 0;
                    shadowDir = cand.direction;
                    shadowTmax = 1e30f;
                }
            } else {
                float uList = (uLight - envProb) / (1.0f - envProb);
                LightConnection cand = sampleLights(s, uList, uPoint, spawn.position);
                f3 wiLight = frame.toLocal(cand.direction);
                bool throughSurface = dot(cand.direction, spawn.normal) < 0.0f;
                if (cand.pdf > 0.0f && wiLight.z != 0.0f && (wiLight.z < 0.0f) == throughSurface) {
                    float neePdf = (1.0f - envProb) * cand.pdf;
                    float weight = !cand.delta && lightSampling == CN_LIGHT_SAMPLING_MIS
                                       ? powerHeuristic(neePdf, bsdfPdf(closure, wiLight, wo))
                                       : 1.0f;
                    unshadowed = (((T * cand.radiance) * evalBsdf(closure, wiLight, wo)) *
                                  std::fabs(wiLight.z)) *
                                 (weight / neePdf);
                    shadowOrigin = throughSurface ? spawn.below : spawn.position;
                    shadowLight = cand.instance;
                    shadowPrim = cand.primitive;
                    shadowDir = cand.direction;
                    shadowTmax = cand.tmax;
                }
            }
            if (any_pos(unshadowed)) {
                haveShadow = true;
                shadowContribution = unshadowed;
            }
        }

        // trace_shadow.slang:27-72: visible iff nothing solid is committed
        // within tmax or the committed hit is the target (instance,
        // primitive). Every fractional-opacity crossing in front of the
        // connection's end that is not the target attenuates it; the
        // product is the order-free fixed-point form above.
        auto traceShadow = [&]() {
            if (!haveShadow) return;
            struct Crossing {
                float t;
                uint32_t inst, prim;
            };
            std::vector<Crossing> list;
            Ray sray{shadowOrigin, shadowDir, shadowTmax};
            HitRecord sh = closest_hit(s, sray, MASK_ALL, [&](uint32_t i, uint32_t p, float t) {
                if (i == shadowLight && p == shadowPrim) return true;
                list.push_back({t, i, p});
                return false;
            });
            if (sh.hit && !(sh.instance == shadowLight && sh.primitive == shadowPrim)) return;
            float tEnd = sh.hit ? sh.t : shadowTmax;
            int64_t logSum = 0;
            for (const Crossing& c : list) {
                if (c.t < tEnd) logSum += s.instances[c.inst].log2_transmittance;
            }
            if (logSum < ATTENUATION_CUTOFF) return;
            float attenuation = exp2_fixed(int32_t(logSum));
            add4(out.radiance, shadowContribution * attenuation, 0.0f);
        };

        // Scatter (shade_surface.slang:453-464).
        float uLobe = sample_1d(pixel, sample, bounce_dimension(bounce, BSDF_LOBE));
        f2 u = sample_2d(pixel, sample, bounce_dimension(bounce, BSDF_DIR));
        BsdfSample scatter = sampleBsdf(closure, wo, uLobe, u);
        f3 wiWorld = frame.toWorld(scatter.wi);
        f3 scattered = splat(0.0f);
        if (scatter.pdf > 0.0f && scatter.wi.z != 0.0f &&
            (scatter.wi.z > 0.0f) == (dot(wiWorld, spawn.normal) > 0.0f)) {
            scattered = (T * evalBsdf(closure, scatter.wi, wo)) *
                        (std::fabs(scatter.wi.z) / scatter.pdf);
        }

        float survival = std::fmax(scattered.x, std::fmax(scattered.y, scattered.z));
        if (survival <= 0.0f) {
            if (aovOn && any_pos(guide)) recordGuides(guide);
            add4(out.radiance, contribution, 1.0f);
            traceShadow();
            return;
        }
        if (bounce >= ROULETTE_START) {
            survival = std::fmin(1.0f, survival);
            if (sample_1d(pixel, sample, bounce_dimension(bounce, RR)) >= survival) {
                if (aovOn && any_pos(guide)) recordGuides(guide);
                add4(out.radiance, contribution, 1.0f);
                traceShadow();
                return;
            }
            scattered = scattered / survival;
        }

        if (aovOn) {
            float record = aovRecordFraction(closure, scatter.lobe);
            if (record > 0.0f && any_pos(guide)) recordGuides(guide * record);
            guideState = (guide * (1.0f - record)) * aovPassTint(closure, scatter.lobe);
        }

        if (scatter.lobe == LOBE_TRANSMISSION && material.thinWalled == 0) {
            medium = medium == hit.instance ? MEDIUM_NONE : hit.instance;
        }
        bool transmitted = scatter.wi.z < 0.0f;
        origin = transmitted ? spawn.below : spawn.position;
        direction = wiWorld;
        throughput = {scattered.x, scattered.y, scattered.z, scatter.pdf};
        state = (scatter.lobe << 24) | (medium & 0xffffff);
        if (any_pos(contribution)) add4(out.radiance, contribution, 0.0f);
        traceShadow();
    }
}

void trace_rows(const orc_scene* s, uint32_t width, uint32_t height, uint32_t row0,
                uint32_t row1, uint32_t sample, uint32_t maxBounces, uint32_t lightSampling,
                bool aov, float* radiance, float* albedo, float* normal, float* depth,
                int threads) {
    float aperture;
    Basis basis = camera_basis(s->camera, float(width) / float(height), &aperture);
    if (threads < 1) threads = 1;
    auto worker = [&](int id) {
        for (uint32_t row = row0 + uint32_t(id); row < row1; row += uint32_t(threads)) {
            for (uint32_t x = 0; x < width; ++x) {
                uint32_t pixel = row * width + x;
                PixelOut o;
                trace_path(*s, basis, aperture, width, height, pixel, sample, maxBounces,
                           lightSampling, aov, o);
                size_t k = size_t(row - row0) * width + x;
                radiance[4 * k + 0] = o.radiance.x;
                radiance[4 * k + 1] = o.radiance.y;
                radiance[4 * k + 2] = o.radiance.z;
                radiance[4 * k + 3] = o.radiance.w;
                if (aov) {
                    if (albedo) {
                        albedo[4 * k + 0] = o.albedo.x;
                        albedo[4 * k + 1] = o.albedo.y;
                        albedo[4 * k + 2] = o.albedo.z;
                        albedo[4 * k + 3] = o.albedo.w;
                    }
                    if (normal) {
                        normal[4 * k + 0] = o.normal.x;
                        normal[4 * k + 1] = o.normal.y;
                        normal[4 * k + 2] = o.normal.z;
                        normal[4 * k + 3] = o.normal.w;
                    }
                    if (depth) depth[k] = o.depth;
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto& th : pool) th.join();
}

// accumulate.slang:36-44
inline void accumulate4(const float* sample, float* sum, size_t k, bool reset) {
    float c[4] = {sample[4 * k], sample[4 * k + 1], sample[4 * k + 2], sample[4 * k + 3]};
    bool bad = false;
    for (float v : c) bad = bad || std::isnan(v) || std::isinf(v);
    if (bad) c[0] = c[1] = c[2] = c[3] = 0.0f;
    for (int i = 0; i < 4; ++i) sum[4 * k + i] = reset ? c[i] : sum[4 * k + i] + c[i];
}

}  // namespace

extern "C" {

void orc_sample_2d(uint32_t pixel, uint32_t sample, uint32_t dimension, float out[2]) {
    f2 p = sample_2d(pixel, sample, dimension);
    out[0] = p.x;
    out[1] = p.y;
}

float orc_sample_1d(uint32_t pixel, uint32_t sample, uint32_t dimension) {
    return sample_1d(pixel, sample, dimension);
}

float orc_hash_uniform(uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t crossing) {
    return hash_uniform(pixel, sample, bounce, crossing);
}

float orc_crossing_uniform(uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t instance, uint32_t primitive) {
    return crossing_uniform(pixel, sample, bounce, instance, primitive);
}
int32_t orc_log2_transmittance_fixed(float opacity) { return log2_transmittance_fixed(opacity); }
float orc_exp2_fixed(int32_t sum) { return exp2_fixed(sum); }

uint32_t orc_lowbias32(uint32_t x) { return lowbias32(x); }
uint32_t orc_laine_karras(uint32_t x, uint32_t seed) { return laine_karras(x, seed); }
uint32_t orc_sobol(uint32_t index, uint32_t axis) { return sobol(index, axis); }
uint32_t orc_reverse_bits(uint32_t x) { return reverse_bits(x); }

void orc_build_lights(const float* tri_corners, const float* tri_emission,
                      const uint32_t* tri_ids, uint32_t tri_count,
                      const cn_delta_light* deltas, uint32_t delta_count,
                      cn_light_record* out, double* total_power) {
    std::vector<TriangleLight> tris(tri_count);
    for (uint32_t i = 0; i < tri_count; ++i) {
        const float* c = tri_corners + 9 * i;
        tris[i].corners[0] = V(c[0], c[1], c[2]);
        tris[i].corners[1] = V(c[3], c[4], c[5]);
        tris[i].corners[2] = V(c[6], c[7], c[8]);
        tris[i].emission = V(tri_emission[3 * i], tri_emission[3 * i + 1], tri_emission[3 * i + 2]);
        tris[i].instance = tri_ids[2 * i];
        tris[i].primitive = tri_ids[2 * i + 1];
    }
    std::vector<LightRec> recs = build_lights(tris, deltas, delta_count, total_power);
    std::memcpy(out, recs.data(), recs.size() * sizeof(LightRec));
}

void orc_env_tables(uint32_t width, uint32_t height, const float* texels, float* marginal,
                    float* conditional, float* pdfs, double* power) {
    EnvTables t = env_tables(width, height, texels);
    std::memcpy(marginal, t.marginal.data(), t.marginal.size() * 4);
    std::memcpy(conditional, t.conditional.data(), t.conditional.size() * 4);
    std::memcpy(pdfs, t.pdfs.data(), t.pdfs.size() * 4);
    *power = t.power;
}

float orc_select_probability(double env_power, double light_power) {
    return select_probability(env_power, light_power);
}

float orc_luminance(const float rgb[3]) { return luminance(V(rgb[0], rgb[1], rgb[2])); }

void orc_acescg_from_rec709(const float rec709[3], float out[3]) {
    f3 c = acescg_from_rec709(V(rec709[0], rec709[1], rec709[2]));
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}

void orc_camera_basis(const cn_camera* camera, float aspect, float out[9]) {
    float aperture;
    Basis b = camera_basis(*camera, aspect, &aperture);
    const f3 v[3] = {b.right, b.up, b.forward};
    for (int i = 0; i < 3; ++i) {
        out[3 * i] = v[i].x;
        out[3 * i + 1] = v[i].y;
        out[3 * i + 2] = v[i].z;
    }
}

orc_scene* orc_scene_create(const cn_scene_desc* desc, const float* bsdf_tables) {
    orc_scene* s = new orc_scene();
    s->camera = desc->camera;
    s->tables.assign(bsdf_tables, bsdf_tables + 16928);
    s->meshes.resize(desc->mesh_count);
    for (uint32_t m = 0; m < desc->mesh_count; ++m) {
        const cn_mesh& src = desc->meshes[m];
        MeshData& d = s->meshes[m];
        d.positions.assign(src.positions, src.positions + 3 * size_t(src.vertex_count));
        d.normals.assign(src.normals, src.normals + 3 * size_t(src.vertex_count));
        if (src.uvs) d.uvs.assign(src.uvs, src.uvs + 2 * size_t(src.vertex_count));
        else d.uvs.assign(2 * size_t(src.vertex_count), 0.0f);
        d.indices.assign(src.indices, src.indices + 3 * size_t(src.triangle_count));
    }
    std::vector<TriangleLight> triLights;
    for (uint32_t i = 0; i < desc->instance_count; ++i) {
        const cn_instance& src = desc->instances[i];
        Instance inst{};
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 4; ++c) inst.o2w[r][c] = src.object_to_world[4 * r + c];
        }
        invert_affine(inst.o2w, inst.w2o);
        inst.mesh = src.mesh;
        inst.mask = src.camera_visible ? MASK_ALL : (MASK_ALL & ~MASK_CAMERA);
        inst.opaque = src.material.opacity >= 1.0f;
        inst.log2_transmittance = inst.opaque ? 0 : log2_transmittance_fixed(src.material.opacity);
        inst.light = LIGHT_NONE;
        Material m = lift_material(src.material);
        s->materials.push_back(m);
        const MeshData& mesh = s->meshes[src.mesh];
        uint32_t ntri = uint32_t(mesh.indices.size() / 3);
        // scene/mod.rs:268-281: an object whose emission is nonzero is a
        // light, one record per triangle in primitive order; its geometry
        // record points at the first (scene/mod.rs:514-519).
        bool emissive = m.emission.x != 0.0f || m.emission.y != 0.0f || m.emission.z != 0.0f;
        if (emissive && ntri > 0) inst.light = uint32_t(triLights.size());
        for (uint32_t p = 0; p < ntri; ++p) {
            f3 w[3];
            for (int k = 0; k < 3; ++k) w[k] = transform_point(inst.o2w, fetch3(mesh.positions, mesh.indices[3 * p + k]));
            s->tris.push_back(WorldTri{w[0], w[1], w[2], i, p});
            if (emissive) triLights.push_back(TriangleLight{{w[0], w[1], w[2]}, m.emission, i, p});
        }
        s->instances.push_back(inst);
    }
    double lightPower = 0.0;
    s->lights = build_lights(triLights, desc->delta_lights, desc->delta_light_count, &lightPower);
    s->env_w = desc->environment.width;
    s->env_h = desc->environment.height;
    s->env_texels.assign(desc->environment.texels,
                         desc->environment.texels + 4 * size_t(s->env_w) * s->env_h);
    s->env = env_tables(s->env_w, s->env_h, desc->environment.texels);
    s->select_prob = select_probability(s->env.power, lightPower);
    build_bvh(*s);
    return s;
}

void orc_scene_destroy(orc_scene* scene) { delete scene; }
uint32_t orc_scene_light_count(orc_scene* scene) { return uint32_t(scene->lights.size()); }
float orc_scene_select_prob(orc_scene* scene) { return scene->select_prob; }
void orc_scene_light_records(orc_scene* scene, cn_light_record* out) {
    std::memcpy(out, scene->lights.data(), scene->lights.size() * sizeof(LightRec));
}

void orc_trace(orc_scene* scene, uint32_t width, uint32_t height, uint32_t sample,
               uint32_t max_bounces, uint32_t light_sampling, int aov, float* radiance,
               float* albedo, float* normal, float* depth, int threads) {
    trace_rows(scene, width, height, 0, height, sample, max_bounces, light_sampling, aov != 0,
               radiance, albedo, normal, depth, threads);
}

void orc_trace_rows(orc_scene* scene, uint32_t width, uint32_t height, uint32_t row0,
                    uint32_t row1, uint32_t sample, uint32_t max_bounces,
                    uint32_t light_sampling, int aov, float* radiance, float* albedo,
                    float* normal, float* depth, int threads) {
    trace_rows(scene, width, height, row0, row1, sample, max_bounces, light_sampling, aov != 0,
               radiance, albedo, normal, depth, threads);
}

void orc_accumulate(orc_scene* scene, uint32_t width, uint32_t height, uint32_t first,
                    uint32_t count, uint32_t max_bounces, uint32_t light_sampling,
                    float* beauty_sum, float* albedo_sum, float* normal_sum, float* depth_sum,
                    int threads) {
    size_t n = size_t(width) * height;
    std::vector<float> rad(4 * n), alb(4 * n), nrm(4 * n), dep(n);
    for (uint32_t k = 0; k < count; ++k) {
        uint32_t sample = first + k;
        trace_rows(scene, width, height, 0, height, sample, max_bounces, light_sampling, true,
                   rad.data(), alb.data(), nrm.data(), dep.data(), threads);
        bool reset = sample == 0;
        for (size_t p = 0; p < n; ++p) {
            accumulate4(rad.data(), beauty_sum, p, reset);
            accumulate4(alb.data(), albedo_sum, p, reset);
            accumulate4(nrm.data(), normal_sum, p, reset);
            float d = dep[p];
            if (std::isnan(d)) d = 0.0f;
            depth_sum[p] = reset ? d : depth_sum[p] + d;
        }
    }
}

void orc_accumulate_rows(orc_scene* scene, uint32_t width, uint32_t height, uint32_t row0, uint32_t row1,
                         uint32_t first, uint32_t count, uint32_t max_bounces, uint32_t light_sampling,
                         float* beauty_sum, int threads) {
    // accumulate.slang:46-67 per pixel of a band of rows: the beauty sum of
    // samples first .. first+count-1, folded in sample order. Threads own
    // pixels (not rows), so narrow bands still use every core.
    float aperture;
    Basis basis = camera_basis(scene->camera, float(width) / float(height), &aperture);
    if (threads < 1) threads = 1;
    const size_t pixels = size_t(row1 - row0) * width;
    auto worker = [&](int id) {
        for (size_t k = size_t(id); k < pixels; k += size_t(threads)) {
            const uint32_t pixel = uint32_t(row0 * width + k);
            for (uint32_t i = 0; i < count; ++i) {
                const uint32_t sample = first + i;
                PixelOut o;
                trace_path(*scene, basis, aperture, width, height, pixel, sample, max_bounces, light_sampling,
                           false, o);
                float c[4] = {o.radiance.x, o.radiance.y, o.radiance.z, o.radiance.w};
                accumulate4(c, beauty_sum + 4 * k, 0, sample == 0);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto& th : pool) th.join();
}

void orc_accumulate_pixels(orc_scene* scene, uint32_t width, uint32_t height, const uint32_t* pixels,
                           uint32_t pixel_count, uint32_t first, uint32_t count, uint32_t max_bounces,
                           uint32_t light_sampling, float* beauty_sum, int threads) {
    // accumulate.slang:46-67 for an arbitrary list of pixel indices of a
    // width x height wave (scattered tiles): sums in sample order, one
    // 4-float entry per listed pixel.
    float aperture;
    Basis basis = camera_basis(scene->camera, float(width) / float(height), &aperture);
    if (threads < 1) threads = 1;
    auto worker = [&](int id) {
        for (size_t k = size_t(id); k < pixel_count; k += size_t(threads)) {
            for (uint32_t i = 0; i < count; ++i) {
                const uint32_t sample = first + i;
                PixelOut o;
                trace_path(*scene, basis, aperture, width, height, pixels[k], sample, max_bounces, light_sampling,
                           false, o);
                float c[4] = {o.radiance.x, o.radiance.y, o.radiance.z, o.radiance.w};
                accumulate4(c, beauty_sum + 4 * k, 0, sample == 0);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto& th : pool) th.join();
}

void orc_first_hits(orc_scene* scene, uint32_t width, uint32_t height, uint32_t sample,
                    uint32_t* instance, uint32_t* primitive, float* bary, int threads) {
    float aperture;
    Basis basis = camera_basis(scene->camera, float(width) / float(height), &aperture);
    f3 camPos = V(scene->camera.position[0], scene->camera.position[1], scene->camera.position[2]);
    if (threads < 1) threads = 1;
    auto worker = [&](int id) {
        for (uint32_t row = uint32_t(id); row < height; row += uint32_t(threads)) {
            for (uint32_t x = 0; x < width; ++x) {
                uint32_t pixel = row * width + x;
                f2 jitter = sample_2d(pixel, sample, CAMERA_JITTER);
                float ndcx = 2.0f * ((float(x) + jitter.x) / float(width)) - 1.0f;
                float ndcy = 1.0f - 2.0f * ((float(row) + jitter.y) / float(height));
                f3 target = (basis.forward + ndcx * basis.right) + ndcy * basis.up;
                f3 origin = camPos;
                f3 direction = normalize(target);
                if (aperture > 0.0f) {
                    f2 u = sample_2d(pixel, sample, CAMERA_LENS);
                    float ox = 2.0f * u.x - 1.0f, oy = 2.0f * u.y - 1.0f;
                    f2 disk{0.0f, 0.0f};
                    if (!(ox == 0.0f && oy == 0.0f)) {
                        float r, theta;
                        if (std::fabs(ox) > std::fabs(oy)) {
                            r = ox;
                            theta = (PI_SLANG / 4.0f) * (oy / ox);
                        } else {
                            r = oy;
                            theta = (PI_SLANG / 2.0f) - (PI_SLANG / 4.0f) * (ox / oy);
                        }
                        disk = {r * std::cos(theta), r * std::sin(theta)};
                    }
                    disk = {aperture * disk.x, aperture * disk.y};
                    f3 offset = disk.x * normalize(basis.right) + disk.y * normalize(basis.up);
                    origin = origin + offset;
                    direction = normalize(target - offset);
                }
                HitRecord h = closest_hit(*scene, Ray{origin, direction, 1e30f}, MASK_CAMERA,
                                          [&](uint32_t inst, uint32_t prim, float) {
                                              return crossing_uniform(pixel, sample, 0, inst, prim) <
                                                     scene->materials[inst].opacity;
                                          });
                instance[pixel] = h.hit ? h.instance : 0xffffffffu;
                primitive[pixel] = h.hit ? h.primitive : 0xffffffffu;
                bary[2 * pixel] = h.hit ? h.b1 : 0.0f;
                bary[2 * pixel + 1] = h.hit ? h.b2 : 0.0f;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto& th : pool) th.join();
}

void orc_env_sample(orc_scene* scene, const float* u, uint32_t count, float* out) {
    for (uint32_t i = 0; i < count; ++i) {
        const f2 ui{u[2 * i], u[2 * i + 1]};
        const CdfSlot row = sampleCdf(scene->env.marginal.data(), scene->env_h, ui.x);
        const CdfSlot col =
            sampleCdf(scene->env.conditional.data() + row.index * (scene->env_w + 1), scene->env_w, ui.y);
        const EnvSample e = env_sample(*scene, ui);
        float* o = out + 12 * size_t(i);
        const float vals[9] = {e.direction.x, e.direction.y, e.direction.z, e.pdf, e.radiance.x,
                               e.radiance.y, e.radiance.z, row.position, col.position};
        std::memcpy(o, vals, sizeof vals);
        std::memcpy(o + 9, &row.index, 4);
        std::memcpy(o + 10, &col.index, 4);
        o[11] = 0.0f;
    }
}

}  // extern "C"
