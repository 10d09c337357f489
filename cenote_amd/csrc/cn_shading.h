// cn_shading.h — the OpenPBR closure, environment and light sampling of
// the surface-shading stage, for gfx950.
//
// Restates crates/cenote/shaders/openpbr.slang, environment.slang and
// lights.slang (citations per function). The BSDF tables are read through a
// pointer the kernel chooses: LDS when the kernel staged them there, HBM
// otherwise — the functions below do not care which.
#pragma once

#include "cn_math.h"
#include "cn_records.h"

namespace cn {

constexpr float PI_K = 3.14159265358979323846f;  // raygen.slang:12, environment.slang:25
constexpr float PI_B = 3.14159265358979f;        // openpbr.slang:40
constexpr float MIN_SPEC_ROUGH = 0.035f;         // openpbr.slang:45

// openpbr.slang:86-95 — offsets into the table blob (crates/cenote/src/tables.rs:51-80).
constexpr uint32_t T_GGX_E = 0;
constexpr uint32_t T_GGX_E_AVG = 1024;
constexpr uint32_t T_GLASS_E = 1056;
constexpr uint32_t T_GLASS_E_AVG = 5152;
constexpr uint32_t T_GLASS_INV_E = 5408;
constexpr uint32_t T_GLASS_INV_E_AVG = 9504;
constexpr uint32_t T_DIELECTRIC_E = 9760;
constexpr uint32_t T_SHEEN_A = 13856;
constexpr uint32_t T_SHEEN_B = 14880;
constexpr uint32_t T_SHEEN_R = 15904;
constexpr uint32_t T_FLOATS = BSDF_TABLE_FLOATS;

// openpbr.slang:289-293
constexpr uint32_t LOBE_DIFFUSE = 0;
constexpr uint32_t LOBE_SPECULAR = 1;
constexpr uint32_t LOBE_TRANSMISSION = 2;
constexpr uint32_t LOBE_COAT = 3;
constexpr uint32_t LOBE_FUZZ = 4;

// The material fields the closure reads, lifted from the 144-byte record.
struct Mat {
    v3 baseColor;
    float baseRoughness;
    v3 emission;
    float metalness;
    float specularWeight, specularRoughness, specularIor, transmissionWeight;
    v3 transmissionColor;
    float transmissionDepth;
    v3 coatColor;
    float coatWeight, coatRoughness, coatIor, coatDarkening, fuzzWeight;
    v3 fuzzColor;
    float fuzzRoughness, opacity;
    uint32_t thinWalled;
};

__device__ __forceinline__ Mat load_material(const cn_material* __restrict__ mats, uint32_t index) {
    // 144 bytes = 9 dwordx4 loads.
    const float4* p = reinterpret_cast<const float4*>(mats + index);
    const float4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4], f = p[5], g = p[6], h = p[7];
    Mat m;
    m.baseColor = V3(a.x, a.y, a.z);
    m.baseRoughness = a.w;
    m.emission = V3(b.x, b.y, b.z);
    m.metalness = b.w;
    m.specularWeight = c.x;
    m.specularRoughness = c.y;
    m.specularIor = c.z;
    m.transmissionWeight = c.w;
    m.transmissionColor = V3(d.x, d.y, d.z);
    m.transmissionDepth = d.w;
    m.coatColor = V3(e.x, e.y, e.z);
    m.coatWeight = e.w;
    m.coatRoughness = f.x;
    m.coatIor = f.y;
    m.coatDarkening = f.z;
    m.fuzzWeight = f.w;
    m.fuzzColor = V3(g.x, g.y, g.z);
    m.fuzzRoughness = g.w;
    m.opacity = h.x;
    m.thinWalled = __float_as_uint(h.y);
    return m;
}

// --- tables (openpbr.slang:97-127) ---------------------------------------------

__device__ __forceinline__ float tab1(const float* t, uint32_t off, uint32_t n, float x) {
    const float fx = saturate(x) * float(n - 1);
    const uint32_t i = min(uint32_t(fx), n - 1);
    const uint32_t j = min(i + 1, n - 1);
    return lerpf(t[off + i], t[off + j], fx - float(i));
}

__device__ __forceinline__ float tab2(const float* t, uint32_t off, uint32_t nx, uint32_t ny, float x, float y) {
    const float fy = saturate(y) * float(ny - 1);
    const uint32_t i = min(uint32_t(fy), ny - 1);
    const uint32_t j = min(i + 1, ny - 1);
    return lerpf(tab1(t, off + nx * i, nx, x), tab1(t, off + nx * j, nx, x), fy - float(i));
}

__device__ __forceinline__ float tab3(const float* t, uint32_t off, float x, float y, float z) {
    const float fz = saturate(z) * 15.0f;
    const uint32_t i = min(uint32_t(fz), 15u);
    const uint32_t j = min(i + 1, 15u);
    return lerpf(tab2(t, off + 256u * i, 16, 16, x, y), tab2(t, off + 256u * j, 16, 16, x, y), fz - float(i));
}

__device__ __forceinline__ float ior_remap(float eta) { return sqrtf(fabsf(eta - 1.0f) / (eta + 1.0f)); }

// --- EON diffuse (openpbr.slang:134-175) ----------------------------------------

__device__ __forceinline__ float fon_a() { return 0.5f - 2.0f / (3.0f * PI_B); }
__device__ __forceinline__ float fon_b() { return 2.0f / 3.0f - 28.0f / (15.0f * PI_B); }

__device__ __forceinline__ float fon_albedo(float mu, float r) {
    const float g = 1.0f - mu;
    const float gOverPi = g * (0.0571085289f + g * (0.491881867f + g * (-0.332181442f + g * 0.0714429953f)));
    return (1.0f + r * gOverPi) / (1.0f + fon_a() * r);
}

__device__ __forceinline__ v3 eon_brdf(v3 rho, float r, v3 wi, v3 wo) {
    const float s = dot(wi, wo) - wi.z * wo.z;
    const float sOverT = s > 0.0f ? s / fmaxf(wi.z, wo.z) : s;
    const float A = 1.0f / (1.0f + fon_a() * r);
    const v3 single = ((rho / PI_B) * A) * (1.0f + r * sOverT);
    const float albedoI = fon_albedo(wi.z, r);
    const float albedoO = fon_albedo(wo.z, r);
    const float albedoAvg = A * (1.0f + fon_b() * r);
    const v3 rhoMulti = ((rho * rho) * albedoAvg) / (S3(1.0f) - rho * (1.0f - albedoAvg));
    const float eps = 1e-7f;
    const v3 multi = (((rhoMulti / PI_B) * fmaxf(eps, 1.0f - albedoO)) * fmaxf(eps, 1.0f - albedoI)) /
                     fmaxf(eps, 1.0f - albedoAvg);
    return single + multi;
}

// --- GGX, Fresnel (openpbr.slang:179-277) -----------------------------------------

__device__ __forceinline__ float ggx_d(float alpha, float hz) {
    const float a2 = alpha * alpha;
    const float denom = hz * hz * (a2 - 1.0f) + 1.0f;
    return hz > 0.0f ? a2 / (PI_B * denom * denom) : 0.0f;
}

__device__ __forceinline__ float smith_g1(float alpha, float vz) {
    const float a2 = alpha * alpha;
    return 2.0f * fabsf(vz) / (fabsf(vz) + sqrtf(vz * vz * (1.0f - a2) + a2));
}

// Dupuy & Benyoub spherical-cap VNDF sampling (openpbr.slang:198-210).
__device__ __forceinline__ v3 sample_vndf(v3 wo, float alpha, f2 u) {
    const v3 woStd = normalize(V3(wo.x * alpha, wo.y * alpha, wo.z));
    const float phi = 2.0f * PI_B * u.x;
    const float z = fmaf(1.0f - u.y, 1.0f + woStd.z, -woStd.z);
    const float sinTheta = sqrtf(clampf(1.0f - z * z, 0.0f, 1.0f));
    float sp, cp;
    sp = sinf(phi);
    cp = cosf(phi);
    const v3 c = V3(sinTheta * cp, sinTheta * sp, z);
    const v3 hStd = c + woStd;
    return normalize(V3(hStd.x * alpha, hStd.y * alpha, hStd.z));
}

__device__ __forceinline__ float vndf_reflect_pdf(float alpha, v3 wo, v3 h) {
    return smith_g1(alpha, wo.z) * ggx_d(alpha, h.z) / (4.0f * wo.z);
}

__device__ __forceinline__ v3 fresnel_schlick(v3 f0, float cosine) {
    const float m = 1.0f - fabsf(cosine);
    const float m2 = m * m;
    return f0 + (((1.0f - f0) * m2) * m2) * m;
}

__device__ __forceinline__ float fresnel_dielectric(float eta, float cosine) {
    const float mu = fabsf(cosine);
    const float sin2 = 1.0f - mu * mu;
    const float cos2t = 1.0f - sin2 / (eta * eta);
    if (cos2t < 0.0f) return 1.0f;
    const float t0 = sqrtf(cos2t);
    const float t1 = eta * t0;
    const float t2 = eta * mu;
    const float rs = (mu - t1) / (mu + t1);
    const float rp = (t0 - t2) / (t0 + t2);
    return (rs * rs + rp * rp) / 2.0f;
}

__device__ __forceinline__ float fresnel_dielectric_avg(float eta) {
    if (eta < 1.0f) return 0.997118f + eta * (0.1014f - eta * (0.965241f + eta * 0.130607f));
    return (eta - 1.0f) / (4.08567f + 1.00071f * eta);
}

__device__ __forceinline__ v3 fresnel_ms(v3 fss, float eAvg) {
    return (fss * eAvg) / (S3(1.0f) - fss * (1.0f - eAvg));
}

// --- the closure (openpbr.slang:295-587) -----------------------------------------

struct Closure {
    Mat m;
    float alpha, coatAlpha, eta;
    bool exiting, hasSpecular;
    float fuzzA, fuzzB, fuzzR, fuzzAtten;
    float coatE, coatScale;
    v3 baseWeight;
    float specE;
    v3 conductorScale;
    float dielectricScale;
    v3 glassScale, glassTint;
    v3 thinR, thinT;
    float thinAlpha, thinScaleR, thinScaleT, thinTShare;
    float wFuzz, wCoat, wSpec, wDiffuse, wTotal;
    v3 aovAlbedo;
};

// openpbr.slang:348-352
__device__ __forceinline__ v3 reflection_scale(const float* t, float rough, float mu, v3 fss) {
    const float e = fmaxf(1e-4f, tab2(t, T_GGX_E, 32, 32, rough, mu));
    const float eAvg = tab1(t, T_GGX_E_AVG, 32, rough);
    return 1.0f + fresnel_ms(fss, eAvg) * ((1.0f - e) / e);
}

// openpbr.slang:357-363
__device__ __forceinline__ float dielectric_albedo(const float* t, float rough, float mu, float eta) {
    if (eta < 1.0f) return fresnel_dielectric(eta, mu);
    return tab3(t, T_DIELECTRIC_E, rough, mu, ior_remap(eta));
}

// openpbr.slang:371-384
__device__ __forceinline__ v3 coat_darkening(const Mat& m, float rough, float mu, float baseFresnel) {
    const float fssCoat = fresnel_dielectric_avg(m.coatIor);
    const float kRough = 1.0f - (1.0f - fssCoat) / (m.coatIor * m.coatIor);
    const float kSmooth = fresnel_dielectric(m.coatIor, mu);
    const float roughDiffuse = lerpf(1.0f, rough, saturate(m.specularWeight * baseFresnel));
    const float baseRough = lerpf(roughDiffuse, rough, m.metalness);
    const float k = lerpf(kSmooth, kRough, baseRough);
    const v3 baseAlbedo = vlerp(m.baseColor, S3(1.0f), (1.0f - m.metalness) * m.transmissionWeight);
    const v3 darkened = (1.0f - k) / vmax(S3(1e-4f), 1.0f - baseAlbedo * k);
    return vlerp(S3(1.0f), darkened, m.coatWeight * m.coatDarkening);
}

// openpbr.slang:391-587
__device__ inline Closure resolve_closure(const float* t, const Mat& m, v3 wo, bool exiting) {
    Closure c;
    c.m = m;
    c.exiting = exiting;
    c.hasSpecular = exiting || m.metalness > 0.0f || m.transmissionWeight > 0.0f || m.specularWeight > 0.0f;
    const float mu = wo.z;
    const bool coated = m.coatWeight > 0.0f && !exiting;

    float rough = clampf(m.specularRoughness, MIN_SPEC_ROUGH, 1.0f);
    const float coatRough = clampf(m.coatRoughness, MIN_SPEC_ROUGH, 1.0f);
    if (coated) {
        const float variance =
            fminf(1.0f, rough * rough * rough * rough + 2.0f * coatRough * coatRough * coatRough * coatRough);
        rough = lerpf(rough, sqrtf(sqrtf(variance)), m.coatWeight);
    }
    c.alpha = rough * rough;
    c.coatAlpha = coatRough * coatRough;

    float etaS = fmaxf(m.specularIor, 1e-4f);
    if (coated) {
        const float etaBc = m.coatIor > etaS ? m.coatIor / etaS : etaS / m.coatIor;
        etaS = lerpf(etaS, etaBc, m.coatWeight);
    }
    const float f0 = (etaS - 1.0f) * (etaS - 1.0f) / ((etaS + 1.0f) * (etaS + 1.0f));
    const float weight = clampf(m.specularWeight, 0.0f, f0 > 1e-6f ? 1.0f / f0 : 0.0f);
    const float epsilon = signf(etaS - 1.0f) * sqrtf(weight * f0);
    float eta = (1.0f + epsilon) / fmaxf(1e-3f, 1.0f - epsilon);
    if (fabsf(eta - 1.0f) < 1e-3f) eta = eta < 1.0f ? 1.0f - 1e-3f : 1.0f + 1e-3f;
    c.eta = exiting ? 1.0f / eta : eta;

    c.fuzzA = 0.0f;
    c.fuzzB = 0.0f;
    c.fuzzR = 0.0f;
    c.fuzzAtten = 1.0f;
    if (m.fuzzWeight > 0.0f && !exiting) {
        const float fr = clampf(m.fuzzRoughness, 1e-3f, 1.0f);
        const float a = tab2(t, T_SHEEN_A, 32, 32, mu, fr);
        const float r = tab2(t, T_SHEEN_R, 32, 32, mu, fr);
        if (fabsf(a) >= 1e-5f && r >= 1e-5f) {
            c.fuzzA = a;
            c.fuzzB = tab2(t, T_SHEEN_B, 32, 32, mu, fr);
            c.fuzzR = r;
            c.fuzzAtten = 1.0f - m.fuzzWeight * r;
        }
    }

    c.coatE = 0.0f;
    c.coatScale = 1.0f;
    v3 coatAtten = S3(1.0f);
    if (coated) {
        c.coatE = dielectric_albedo(t, coatRough, mu, m.coatIor);
        const float fssCoat = fresnel_dielectric_avg(m.coatIor);
        c.coatScale = reflection_scale(t, coatRough, mu, S3(fssCoat)).x;
        coatAtten = vlerp(S3(1.0f), m.coatColor * (1.0f - c.coatE), m.coatWeight) * coat_darkening(m, rough, mu, f0);
    }
    c.baseWeight = S3(c.fuzzAtten) * coatAtten;

    c.specE = c.hasSpecular ? dielectric_albedo(t, rough, mu, c.eta) : 0.0f;
    c.conductorScale = reflection_scale(t, rough, mu, vlerp(m.baseColor, S3(1.0f), 1.0f / 21.0f));
    c.dielectricScale = reflection_scale(t, rough, mu, S3(fresnel_dielectric_avg(c.eta))).x;

    c.glassTint = m.transmissionDepth == 0.0f ? m.transmissionColor : S3(1.0f);
    float glassEta = c.eta;
    uint32_t glassOff = T_GLASS_E, glassAvgOff = T_GLASS_E_AVG;
    if (glassEta < 1.0f) {
        glassEta = 1.0f / glassEta;
        glassOff = T_GLASS_INV_E;
        glassAvgOff = T_GLASS_INV_E_AVG;
    }
    const float glassZ = ior_remap(glassEta);
    const float glassE = fmaxf(1e-4f, tab3(t, glassOff, rough, mu, glassZ));
    const float glassEAvg = tab2(t, glassAvgOff, 16, 16, rough, glassZ);
    c.glassScale = 1.0f + fresnel_ms(c.glassTint, glassEAvg) * ((1.0f - glassE) / glassE);

    c.thinR = S3(0.0f);
    c.thinT = S3(0.0f);
    c.thinAlpha = c.alpha;
    c.thinScaleR = 1.0f;
    c.thinScaleT = 1.0f;
    c.thinTShare = 0.0f;
    const bool thin = m.thinWalled != 0 && m.transmissionWeight > 0.0f;
    if (thin) {
        const float thinEta = fmaxf(eta, 1.0f / eta);
        const float r1 = fresnel_dielectric(thinEta, mu);
        const float t1 = 1.0f - r1;
        const float cos2t = fmaxf(0.0f, 1.0f - (1.0f - mu * mu) / (thinEta * thinEta));
        v3 absorb = S3(0.0f);
        if (cos2t > 0.0f) {
            const v3 tc = vclamp(m.transmissionColor, 1e-4f, 1.0f);
            const float p = 1.0f / sqrtf(cos2t);
            absorb = V3(powf(tc.x, p), powf(tc.y, p), powf(tc.z, p));
        }
        const v3 denom = vmax(S3(1e-4f), 1.0f - (absorb * absorb) * (r1 * r1));
        c.thinT = (absorb * (t1 * t1)) / denom;
        c.thinR = r1 + c.thinT * (r1 * absorb);
        const float thinRough = clampf(
            rough * sqrtf(3.4f * (thinEta - 1.0f) * (thinEta - 0.5f) * (thinEta - 0.5f) / (thinEta * thinEta * thinEta)),
            MIN_SPEC_ROUGH, 1.0f);
        c.thinAlpha = thinRough * thinRough;
        c.thinScaleR = reflection_scale(t, rough, mu, S3(1.0f)).x;
        c.thinScaleT = reflection_scale(t, thinRough, mu, S3(1.0f)).x;
        const float total = mean3(c.thinR) + mean3(c.thinT);
        c.thinTShare = total > 0.0f ? mean3(c.thinT) / total : 0.0f;
    }

    const float schlickS = saturate((c.specE - f0) / fmaxf(1e-4f, 1.0f - f0));
    const v3 metalE = vlerp(m.baseColor, S3(1.0f), schlickS);
    const float glassTotal = thin ? mean3(c.thinR + c.thinT) : c.specE + (1.0f - c.specE) * mean3(c.glassTint);
    const float dielectricE = lerpf(c.specE, glassTotal, m.transmissionWeight);
    const float baseAvg = mean3(c.baseWeight);
    c.wFuzz = exiting ? 0.0f : m.fuzzWeight * c.fuzzR * mean3(m.fuzzColor);
    c.wCoat = exiting ? 0.0f : c.fuzzAtten * m.coatWeight * c.coatE;
    c.wSpec = c.hasSpecular ? baseAvg * lerpf(dielectricE, mean3(metalE), m.metalness) : 0.0f;
    c.wDiffuse = exiting ? 0.0f
                         : baseAvg * (1.0f - m.metalness) * (1.0f - m.transmissionWeight) * (1.0f - c.specE) *
                               mean3(m.baseColor);
    if (exiting) c.wSpec = 1.0f;
    c.wTotal = c.wFuzz + c.wCoat + c.wSpec + c.wDiffuse;

    const v3 glassColor = thin ? c.thinR + c.thinT : S3(c.specE) + (1.0f - c.specE) * c.glassTint;
    if (exiting) {
        c.aovAlbedo = glassColor;
    } else {
        const v3 dielectricColor = vlerp(S3(c.specE), glassColor, m.transmissionWeight);
        const v3 specColor = c.hasSpecular ? vlerp(dielectricColor, metalE, m.metalness) : S3(0.0f);
        c.aovAlbedo = ((m.fuzzWeight * c.fuzzR) * m.fuzzColor + S3(c.fuzzAtten * m.coatWeight * c.coatE)) +
                      c.baseWeight *
                          (specColor +
                           (((1.0f - m.metalness) * (1.0f - m.transmissionWeight)) * (1.0f - c.specE)) * m.baseColor);
    }
    return c;
}

// openpbr.slang:596-608
__device__ __forceinline__ float aov_record_fraction(const Closure& c, uint32_t lobe) {
    // The three candidates are read unconditionally and pinned in registers:
    // a branch per field lets the optimiser merge the loads into one load at
    // a lobe-dependent offset, which forces the whole Closure into scratch.
    float coat = c.coatAlpha, spec = c.alpha, thin = c.thinAlpha;
    asm("" : "+v"(coat), "+v"(spec), "+v"(thin));
    float alpha;
    if (lobe == LOBE_COAT) alpha = coat;
    else if (lobe == LOBE_SPECULAR) alpha = spec;
    else if (lobe == LOBE_TRANSMISSION) alpha = c.m.thinWalled != 0 ? thin : spec;
    else return 1.0f;
    return saturate(sqrtf(alpha) / 0.15f);
}

// openpbr.slang:614-622
__device__ __forceinline__ v3 aov_pass_tint(const Closure& c, uint32_t lobe) {
    if (lobe == LOBE_TRANSMISSION) return c.glassTint;
    if (lobe == LOBE_SPECULAR) return vlerp(S3(1.0f), c.m.baseColor, c.m.metalness);
    return S3(1.0f);
}

// openpbr.slang:629-647
__device__ __forceinline__ float sheen_density(const Closure& c, v3 w) {
    const float a = c.fuzzA, b = c.fuzzB;
    const float lenSqr = (a * w.x + b * w.z) * (a * w.x + b * w.z) + (a * w.y) * (a * w.y) + w.z * w.z;
    const float ratio = a / lenSqr;
    return fmaxf(w.z, 0.0f) * ratio * ratio / PI_B;
}

__device__ __forceinline__ v3 align_to_view(v3 v, v3 wo, bool inverse) {
    const float len = sqrtf(wo.x * wo.x + wo.y * wo.y);
    if (len < 1e-6f) return v;
    const float cosPhi = wo.x / len;
    const float sinPhi = (inverse ? -wo.y : wo.y) / len;
    return V3(cosPhi * v.x + sinPhi * v.y, cosPhi * v.y - sinPhi * v.x, v.z);
}

// openpbr.slang:650-653
__device__ __forceinline__ float reflection_core(float alpha, v3 wi, v3 wo, v3 h) {
    return ggx_d(alpha, h.z) * smith_g1(alpha, wo.z) * smith_g1(alpha, wi.z) / (4.0f * wo.z * wi.z);
}

// openpbr.slang:660-677
__device__ __forceinline__ v3 base_reflection(const Closure& c, v3 wi, v3 wo, v3 h) {
    const float core = reflection_core(c.alpha, wi, wo, h);
    const float cosVH = dot(h, wo);
    const float fd = fresnel_dielectric(c.eta, cosVH);
    if (c.exiting) return S3(core * fd) * c.glassScale;
    const v3 conductor = fresnel_schlick(c.m.baseColor, cosVH) * c.conductorScale;
    v3 dielectric = c.m.thinWalled != 0
                        ? vlerp(S3(fd * c.dielectricScale), c.thinR * c.thinScaleR, c.m.transmissionWeight)
                        : vlerp(S3(fd * c.dielectricScale), fd * c.glassScale, c.m.transmissionWeight);
    return core * vlerp(dielectric, conductor, c.m.metalness);
}

// openpbr.slang:681-744 — f(wi, wo) without the cosine.
__device__ inline v3 eval_bsdf(const Closure& c, v3 wi, v3 wo) {
    if (wo.z <= 0.0f || wi.z == 0.0f) return S3(0.0f);
    const Mat& m = c.m;
    if (wi.z < 0.0f) {
        const float glassShare = c.exiting ? 1.0f : (1.0f - m.metalness) * m.transmissionWeight;
        if (glassShare <= 0.0f || (c.exiting && m.thinWalled != 0)) return S3(0.0f);
        if (m.thinWalled != 0) {
            const v3 flipped = V3(wi.x, wi.y, -wi.z);
            const v3 h = normalize(flipped + wo);
            const float core = reflection_core(c.thinAlpha, flipped, wo, h);
            return ((c.baseWeight * glassShare) * c.thinT) * c.thinScaleT * core;
        }
        const v3 hU = wo + c.eta * wi;
        const float len2 = dot(hU, hU);
        if (len2 < 1e-9f) return S3(0.0f);
        v3 h = hU / sqrtf(len2);
        h = h.z >= 0.0f ? h : -h;
        const float cosVH = dot(h, wo);
        const float cosLH = dot(h, wi);
        if (cosVH <= 0.0f || cosLH >= 0.0f) return S3(0.0f);
        const float f = fresnel_dielectric(c.eta, cosVH);
        const float core = ggx_d(c.alpha, h.z) * smith_g1(c.alpha, wo.z) * smith_g1(c.alpha, wi.z) * cosVH *
                           (c.eta * c.eta) * fabsf(cosLH) / (wo.z * fabsf(wi.z) * len2);
        return ((((c.baseWeight * glassShare) * (1.0f - f)) * c.glassTint) * c.glassScale) * core;
    }
    v3 result = S3(0.0f);
    if (c.fuzzR > 0.0f) {
        const float density = sheen_density(c, align_to_view(wi, wo, false));
        result = result + ((m.fuzzWeight * m.fuzzColor) * c.fuzzR) * density / wi.z;
    }
    const v3 h = normalize(wi + wo);
    if (m.coatWeight > 0.0f && !c.exiting) {
        const float coatF = fresnel_dielectric(m.coatIor, dot(h, wo));
        result = result + S3(c.fuzzAtten * m.coatWeight * coatF * c.coatScale * reflection_core(c.coatAlpha, wi, wo, h));
    }
    v3 base = S3(0.0f);
    if (c.hasSpecular) base = base_reflection(c, wi, wo, h);
    if (!c.exiting) {
        base = base + (((1.0f - m.metalness) * (1.0f - m.transmissionWeight) * (1.0f - c.specE)) *
                       eon_brdf(m.baseColor, m.baseRoughness, wi, wo));
    }
    return result + c.baseWeight * base;
}

// openpbr.slang:748-752
__device__ __forceinline__ float reflect_probability(const Closure& c, float fresnel) {
    const float glassShare = c.exiting ? 1.0f : (1.0f - c.m.metalness) * c.m.transmissionWeight;
    const float transmit = c.m.thinWalled != 0 ? c.thinTShare : 1.0f - fresnel;
    return 1.0f - glassShare * transmit;
}

// openpbr.slang:756-808
__device__ inline float bsdf_pdf(const Closure& c, v3 wi, v3 wo) {
    if (wo.z <= 0.0f || wi.z == 0.0f || c.wTotal <= 0.0f) return 0.0f;
    if (wi.z < 0.0f) {
        if (c.wSpec <= 0.0f) return 0.0f;
        float pdf;
        if (c.m.thinWalled != 0) {
            const v3 flipped = V3(wi.x, wi.y, -wi.z);
            const v3 h = normalize(flipped + wo);
            pdf = vndf_reflect_pdf(c.thinAlpha, wo, h) * (1.0f - reflect_probability(c, 0.0f));
        } else {
            const v3 hU = wo + c.eta * wi;
            const float len2 = dot(hU, hU);
            if (len2 < 1e-9f) return 0.0f;
            v3 h = hU / sqrtf(len2);
            h = h.z >= 0.0f ? h : -h;
            const float cosVH = dot(h, wo);
            const float cosLH = dot(h, wi);
            if (cosVH <= 0.0f || cosLH >= 0.0f) return 0.0f;
            const float f = fresnel_dielectric(c.eta, cosVH);
            const float vndf = smith_g1(c.alpha, wo.z) * ggx_d(c.alpha, h.z) * cosVH / wo.z;
            const float jacobian = (c.eta * c.eta) * fabsf(cosLH) / len2;
            pdf = vndf * jacobian * (1.0f - reflect_probability(c, f));
        }
        return c.wSpec * pdf / c.wTotal;
    }
    const v3 h = normalize(wi + wo);
    float pdf = 0.0f;
    if (c.wFuzz > 0.0f) pdf += c.wFuzz * sheen_density(c, align_to_view(wi, wo, false));
    if (c.wCoat > 0.0f) pdf += c.wCoat * vndf_reflect_pdf(c.coatAlpha, wo, h);
    if (c.wSpec > 0.0f) {
        const float f = fresnel_dielectric(c.eta, dot(h, wo));
        pdf += c.wSpec * vndf_reflect_pdf(c.alpha, wo, h) * reflect_probability(c, f);
    }
    if (c.wDiffuse > 0.0f) pdf += c.wDiffuse * wi.z / PI_B;
    return pdf / c.wTotal;
}

// f(wi, wo) and pdf(wi | wo) together: eval_bsdf's and bsdf_pdf's values
// bit for bit, with the terms both compute for a reflected direction (the
// half vector, each lobe's D and G1(wo), the dielectric Fresnel, the sheen
// density) computed once, under the union of the two functions' conditions
// and by the same expressions (products reordered only where IEEE
// multiplication commutes); for a transmitted direction, the refraction
// half vector, Fresnel, D and G1(wo) likewise.
struct BsdfEval {
    v3 f;
    float pdf;
};

__device__ inline BsdfEval eval_with_pdf(const Closure& c, v3 wi, v3 wo) {
    BsdfEval r{S3(0.0f), 0.0f};
    if (wo.z <= 0.0f || wi.z == 0.0f) return r;
    const Mat& m = c.m;
    const bool pdfOn = !(c.wTotal <= 0.0f);  // bsdf_pdf's own early exit
    if (wi.z < 0.0f) {  // both functions' transmission branches
        const float glassShare = c.exiting ? 1.0f : (1.0f - m.metalness) * m.transmissionWeight;
        const bool fOn = !(glassShare <= 0.0f || (c.exiting && m.thinWalled != 0));
        const bool pOn = pdfOn && !(c.wSpec <= 0.0f);
        if (!fOn && !pOn) return r;
        if (m.thinWalled != 0) {
            const v3 flipped = V3(wi.x, wi.y, -wi.z);
            const v3 h = normalize(flipped + wo);
            const float D = ggx_d(c.thinAlpha, h.z), G1o = smith_g1(c.thinAlpha, wo.z);
            if (fOn) {
                const float core = D * G1o * smith_g1(c.thinAlpha, flipped.z) / (4.0f * wo.z * flipped.z);
                r.f = ((c.baseWeight * glassShare) * c.thinT) * c.thinScaleT * core;
            }
            if (pOn) {
                const float pdf = G1o * D / (4.0f * wo.z) * (1.0f - reflect_probability(c, 0.0f));
                r.pdf = c.wSpec * pdf / c.wTotal;
            }
            return r;
        }
        const v3 hU = wo + c.eta * wi;
        const float len2 = dot(hU, hU);
        if (len2 < 1e-9f) return r;
        v3 h = hU / sqrtf(len2);
        h = h.z >= 0.0f ? h : -h;
        const float cosVH = dot(h, wo);
        const float cosLH = dot(h, wi);
        if (cosVH <= 0.0f || cosLH >= 0.0f) return r;
        const float f = fresnel_dielectric(c.eta, cosVH);
        const float D = ggx_d(c.alpha, h.z), G1o = smith_g1(c.alpha, wo.z);
        if (fOn) {
            const float core = D * G1o * smith_g1(c.alpha, wi.z) * cosVH * (c.eta * c.eta) * fabsf(cosLH) /
                               (wo.z * fabsf(wi.z) * len2);
            r.f = ((((c.baseWeight * glassShare) * (1.0f - f)) * c.glassTint) * c.glassScale) * core;
        }
        if (pOn) {
            const float vndf = G1o * D * cosVH / wo.z;
            const float jacobian = (c.eta * c.eta) * fabsf(cosLH) / len2;
            const float pdf = vndf * jacobian * (1.0f - reflect_probability(c, f));
            r.pdf = c.wSpec * pdf / c.wTotal;
        }
        return r;
    }
    const bool fuzzF = c.fuzzR > 0.0f, fuzzP = pdfOn && c.wFuzz > 0.0f;
    const bool coatF = m.coatWeight > 0.0f && !c.exiting, coatP = pdfOn && c.wCoat > 0.0f;
    const bool specF = c.hasSpecular, specP = pdfOn && c.wSpec > 0.0f;
    const float density = (fuzzF || fuzzP) ? sheen_density(c, align_to_view(wi, wo, false)) : 0.0f;
    const v3 h = normalize(wi + wo);
    const float cosVH = dot(h, wo);
    float coatD = 0.0f, coatG1o = 0.0f, specD = 0.0f, specG1o = 0.0f, fd = 0.0f;
    if (coatF || coatP) {
        coatD = ggx_d(c.coatAlpha, h.z);
        coatG1o = smith_g1(c.coatAlpha, wo.z);
    }
    if (specF || specP) {
        specD = ggx_d(c.alpha, h.z);
        specG1o = smith_g1(c.alpha, wo.z);
        fd = fresnel_dielectric(c.eta, cosVH);
    }

    // eval_bsdf's reflection branch
    v3 result = S3(0.0f);
    if (fuzzF) result = result + ((m.fuzzWeight * m.fuzzColor) * c.fuzzR) * density / wi.z;
    if (coatF) {
        const float coatFresnel = fresnel_dielectric(m.coatIor, cosVH);
        const float core = coatD * coatG1o * smith_g1(c.coatAlpha, wi.z) / (4.0f * wo.z * wi.z);
        result = result + S3(c.fuzzAtten * m.coatWeight * coatFresnel * c.coatScale * core);
    }
    v3 base = S3(0.0f);
    if (specF) {  // base_reflection
        const float core = specD * specG1o * smith_g1(c.alpha, wi.z) / (4.0f * wo.z * wi.z);
        if (c.exiting) {
            base = S3(core * fd) * c.glassScale;
        } else {
            const v3 conductor = fresnel_schlick(m.baseColor, cosVH) * c.conductorScale;
            v3 dielectric = m.thinWalled != 0
                                ? vlerp(S3(fd * c.dielectricScale), c.thinR * c.thinScaleR, m.transmissionWeight)
                                : vlerp(S3(fd * c.dielectricScale), fd * c.glassScale, m.transmissionWeight);
            base = core * vlerp(dielectric, conductor, m.metalness);
        }
    }
    if (!c.exiting) {
        base = base + (((1.0f - m.metalness) * (1.0f - m.transmissionWeight) * (1.0f - c.specE)) *
                       eon_brdf(m.baseColor, m.baseRoughness, wi, wo));
    }
    r.f = result + c.baseWeight * base;

    // bsdf_pdf's reflection branch
    if (pdfOn) {
        float pdf = 0.0f;
        if (c.wFuzz > 0.0f) pdf += c.wFuzz * density;
        if (c.wCoat > 0.0f) pdf += c.wCoat * (coatG1o * coatD / (4.0f * wo.z));
        if (c.wSpec > 0.0f) pdf += c.wSpec * (specG1o * specD / (4.0f * wo.z)) * reflect_probability(c, fd);
        if (c.wDiffuse > 0.0f) pdf += c.wDiffuse * wi.z / PI_B;
        r.pdf = pdf / c.wTotal;
    }
    return r;
}

struct BsdfSample {
    v3 wi;
    float pdf;
    uint32_t lobe;
};

// openpbr.slang:821-884, up to the pdf: draws the scattered direction and
// lobe into `r` (pdf left 0) and returns whether one was drawn; the pdf of a
// drawn direction is bsdf_pdf(c, r.wi, wo), which eval_with_pdf also gives.
__device__ inline bool sample_bsdf_direction(const Closure& c, v3 wo, float uLobe, f2 u, BsdfSample& r) {
    r = BsdfSample{V3(0.0f, 0.0f, 1.0f), 0.0f, LOBE_DIFFUSE};
    if (c.wTotal <= 0.0f) return false;
    const float pick = uLobe * c.wTotal;
    v3 wi;
    if (pick < c.wFuzz) {
        r.lobe = LOBE_FUZZ;
        const float rr = sqrtf(u.x);
        const float phi = 2.0f * PI_B * u.y;
        const float dx = rr * cosf(phi), dy = rr * sinf(phi);
        const float diskZ = sqrtf(fmaxf(0.0f, 1.0f - (dx * dx + dy * dy)));
        const v3 wiStd = normalize(V3(dx - diskZ * c.fuzzB, dy, diskZ * c.fuzzA));
        wi = align_to_view(wiStd, wo, true);
    } else if (pick < c.wFuzz + c.wCoat) {
        r.lobe = LOBE_COAT;
        const v3 h = sample_vndf(wo, c.coatAlpha, u);
        wi = (2.0f * dot(wo, h)) * h - wo;
    } else if (pick < c.wFuzz + c.wCoat + c.wSpec) {
        const float uBranch = (pick - c.wFuzz - c.wCoat) / fmaxf(1e-6f, c.wSpec);
        if (c.m.thinWalled != 0) {
            if (uBranch < reflect_probability(c, 0.0f)) {
                r.lobe = LOBE_SPECULAR;
                const v3 h = sample_vndf(wo, c.alpha, u);
                wi = (2.0f * dot(wo, h)) * h - wo;
            } else {
                r.lobe = LOBE_TRANSMISSION;
                const v3 h = sample_vndf(wo, c.thinAlpha, u);
                const v3 refl = (2.0f * dot(wo, h)) * h - wo;
                wi = V3(refl.x, refl.y, -refl.z);
            }
        } else {
            const v3 h = sample_vndf(wo, c.alpha, u);
            const float f = fresnel_dielectric(c.eta, dot(h, wo));
            if (uBranch < reflect_probability(c, f)) {
                r.lobe = LOBE_SPECULAR;
                wi = (2.0f * dot(wo, h)) * h - wo;
            } else {
                r.lobe = LOBE_TRANSMISSION;
                // refractDirection, openpbr.slang:270-277
                const float cc = dot(wo, h);
                const float cos2t = 1.0f - (1.0f - cc * cc) / (c.eta * c.eta);
                if (cos2t <= 0.0f) return false;
                wi = h * (cc / c.eta - sqrtf(cos2t)) - wo / c.eta;
            }
        }
    } else {
        r.lobe = LOBE_DIFFUSE;
        // sampleCosineHemisphere, openpbr.slang:170-175
        const float rr = sqrtf(u.x);
        const float phi = 2.0f * PI_B * u.y;
        wi = V3(rr * cosf(phi), rr * sinf(phi), sqrtf(1.0f - u.x));
    }
    r.wi = wi;
    return true;
}

// Duff et al. orthonormal basis (openpbr.slang:888-909).
struct Frame {
    v3 t, b, n;
};
__device__ __forceinline__ Frame make_frame(v3 n) {
    const float s = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (s + n.z);
    const float m = n.x * n.y * a;
    return {V3(1.0f + s * n.x * n.x * a, s * m, -s * n.x), V3(m, s + n.y * n.y * a, -n.y), n};
}
__device__ __forceinline__ v3 to_local(const Frame& f, v3 v) { return V3(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }
__device__ __forceinline__ v3 to_world(const Frame& f, v3 v) { return (v.x * f.t + v.y * f.b) + v.z * f.n; }

// --- environment (environment.slang) ----------------------------------------------

// Software bilinear filter with the reference sampler's address modes
// (REPEAT in u, CLAMP in v; gpu/image.rs:95-103, :263-270). CDNA has no
// sampler path for a raw buffer, so this is four texel loads and exact f32
// weights.
__device__ __forceinline__ v3 env_bilinear(const SceneView& s, float u, float v) {
    const int w = int(s.env_width), h = int(s.env_height);
    const float x = u * float(w) - 0.5f;
    const float y = v * float(h) - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    int x0 = int(fx), y0 = int(fy);
    int x1 = x0 + 1, y1 = y0 + 1;
    // For u in [0, 1], x0 lies in [-1, w - 1] and x1 in [0, w]: one add or
    // subtract wraps them exactly as ((x % w) + w) % w does, without its two
    // integer divisions; any other value (a non-finite u) takes that form.
    if (x0 >= -1 && x0 < w) {
        x0 = x0 < 0 ? x0 + w : x0;
        x1 = x1 >= w ? x1 - w : x1;
    } else {
        x0 = ((x0 % w) + w) % w;
        x1 = ((x1 % w) + w) % w;
    }
    y0 = min(max(y0, 0), h - 1);
    y1 = min(max(y1, 0), h - 1);
    const v3 t00 = xyz(s.env_texels[y0 * w + x0]);
    const v3 t10 = xyz(s.env_texels[y0 * w + x1]);
    const v3 t01 = xyz(s.env_texels[y1 * w + x0]);
    const v3 t11 = xyz(s.env_texels[y1 * w + x1]);
    const v3 top = t00 * (1.0f - ax) + t10 * ax;
    const v3 bottom = t01 * (1.0f - ax) + t11 * ax;
    return top * (1.0f - ay) + bottom * ay;
}

// environment.slang:74-78
__device__ __forceinline__ f2 uv_from_direction(v3 d) {
    return {0.5f + atan2f(d.x, -d.z) / (2.0f * PI_K), acosf(clampf(d.y, -1.0f, 1.0f)) / PI_K};
}

__device__ __forceinline__ v3 env_radiance(const SceneView& s, v3 d) {
    const f2 uv = uv_from_direction(d);
    return env_bilinear(s, uv.x, uv.y);
}

// environment.slang:41-58. The index is the last i in [0, count) with
// cdf[i] <= u; the guide (buckets + 1 entries, buckets a power of two) holds
// that index at u = g / buckets, so for u in bucket g it lies in
// [guide[g], guide[g + 1]] and the search covers only that range — the same
// index the search over the whole table finds, in a few steps instead of
// log2(count) dependent loads.
__device__ __forceinline__ void sample_cdf(const float* __restrict__ cdf, const uint32_t* __restrict__ guide,
                                           uint32_t buckets, uint32_t count, float u, uint32_t* index,
                                           float* position) {
    const uint32_t g = min(uint32_t(u * float(buckets)), buckets - 1);
    uint32_t lo = guide[g], hi = guide[g + 1] + 1;
    while (lo + 1 < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid;
    }
    const float width = cdf[lo + 1] - cdf[lo];
    const float frac = width > 0.0f ? (u - cdf[lo]) / width : 0.0f;
    *index = lo;
    *position = (float(lo) + frac) / float(count);
}

struct EnvSample {
    v3 direction, radiance;
    float pdf;
};

// environment.slang:80-85, :94-106
__device__ inline EnvSample env_sample(const SceneView& s, f2 u) {
    uint32_t row, col;
    float v, uu;
    sample_cdf(s.env_marginal, s.env_marginal_guide, s.env_marginal_buckets, s.env_height, u.x, &row, &v);
    sample_cdf(s.env_conditional + row * (s.env_width + 1),
               s.env_conditional_guide + row * (s.env_conditional_buckets + 1), s.env_conditional_buckets,
               s.env_width, u.y, &col, &uu);
    const float sinTheta = sinf(PI_K * v);
    const float phi = 2.0f * PI_K * (uu - 0.5f);
    const float theta = PI_K * v;
    const float st = sinf(theta);
    EnvSample e;
    e.direction = V3(st * sinf(phi), cosf(theta), -st * cosf(phi));
    e.radiance = env_bilinear(s, uu, v);
    e.pdf = sinTheta > 0.0f ? s.env_pdfs[row * s.env_width + col] / (2.0f * PI_K * PI_K * sinTheta) : 0.0f;
    return e;
}

// environment.slang:110-119
__device__ __forceinline__ float env_pdf(const SceneView& s, v3 d) {
    const float sinTheta = sqrtf(fmaxf(0.0f, 1.0f - d.y * d.y));
    if (sinTheta <= 0.0f) return 0.0f;
    const f2 uv = uv_from_direction(d);
    const uint32_t col = min(uint32_t(uv.x * float(s.env_width)), s.env_width - 1);
    const uint32_t row = min(uint32_t(uv.y * float(s.env_height)), s.env_height - 1);
    return s.env_pdfs[row * s.env_width + col] / (2.0f * PI_K * PI_K * sinTheta);
}

// lights.slang:42-44
__device__ __forceinline__ float power_heuristic(float p, float q) { return (p * p) / (p * p + q * q); }

struct LightConnection {
    v3 direction;
    float tmax;
    v3 radiance;
    float pdf;
    bool delta;
    uint32_t instance, primitive;
};

// lights.slang:93-160: alias-table pick, then a point on the light.
__device__ inline LightConnection sample_lights(const SceneView& s, float u, f2 uPoint, v3 from) {
    const uint32_t count = s.light_count;
    const float scaled = u * float(count);
    const uint32_t slot = min(uint32_t(scaled), count - 1);
    const float4* rec = reinterpret_cast<const float4*>(s.lights + slot);
    float4 r1 = rec[0];  // anchor, pdf
    float4 r2 = rec[1];  // edge1, alias threshold
    float4 r3 = rec[2];  // edge2, alias index
    if (scaled - float(slot) >= r2.w) {
        rec = reinterpret_cast<const float4*>(s.lights + __float_as_uint(r3.w));
        r1 = rec[0];
        r2 = rec[1];
        r3 = rec[2];
    }
    const float4 r4 = rec[3];  // emission, instance
    const uint4 r5 = reinterpret_cast<const uint4*>(rec)[4];  // kind, primitive
    const v3 anchor = xyz(r1), edge1 = xyz(r2), edge2 = xyz(r3);
    const float lpdf = r1.w;
    const uint32_t kind = r5.x;

    LightConnection c;
    c.pdf = 0.0f;
    c.direction = S3(0.0f);
    c.tmax = 1e30f;
    c.radiance = xyz(r4);
    c.delta = kind != LIGHT_KIND_TRIANGLE;
    c.instance = __float_as_uint(r4.w);
    c.primitive = r5.y;
    if (kind == LIGHT_KIND_TRIANGLE) {
        const float su = sqrtf(uPoint.x);
        const float b1 = 1.0f - su, b2 = uPoint.y * su;
        const v3 position = (anchor + b1 * edge1) + b2 * edge2;
        const v3 delta = position - from;
        const float r2s = dot(delta, delta);
        const v3 normal = normalize(cross(edge1, edge2));
        const v3 direction = delta * rsqrt_exact(r2s);
        const float cosLight = -dot(normal, direction);
        if (r2s > 0.0f && cosLight > 0.0f && lpdf > 0.0f) {
            c.direction = direction;
            c.pdf = lpdf * r2s / cosLight;
            c.tmax = sqrtf(r2s) * 1.001f;
        }
    } else if (kind == LIGHT_KIND_DISTANT) {
        c.direction = -anchor;
        c.pdf = lpdf;
    } else {
        const v3 delta = anchor - from;
        const float r2s = dot(delta, delta);
        if (r2s > 0.0f && lpdf > 0.0f) {
            const float r = sqrtf(r2s);
            c.direction = delta / r;
            c.radiance = c.radiance / r2s;
            c.pdf = lpdf;
            c.tmax = r;
        }
    }
    return c;
}

}  // namespace cn
