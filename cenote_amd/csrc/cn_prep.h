// cn_prep.h — host-side scene preparation: the light table, the environment
// sampling tables, transforms and the camera basis. Everything here runs
// once per scene on the host, in the operation order the reference's host
// code uses, because its outputs (alias tables, CDFs) must match bit for bit.
#pragma once

#include <stdint.h>

#include <vector>

#include "../../include/cenote_hip.h"
#include "cn_math.h"
#include "cn_records.h"

namespace cn {

// An emissive triangle in world space (lights.rs:39-52).
struct EmissiveTriangle {
    v3 corners[3];
    v3 emission;
    uint32_t instance, primitive;
};

// color.rs:29-31
float luminance(v3 c);

// A material's lobe set as a class in 0..31 (metallic, dielectric base,
// transmission, coat, fuzz): materials of one class take the same branches
// through the closure.
uint32_t material_class(const cn_material& m);

// A fractional-opacity instance's log2(1 - opacity) in 2^-24 units, rounded
// to nearest and floored at LOG2_FIXED_FLOOR (one crossing that dark
// already ends a connection): the per-crossing term of the shadow
// attenuation's order-free product (exp2_fixed, cn_math.h).
int32_t log2_transmittance_fixed(float opacity);

// lights.rs:175-224, :229-231. Returns the records; *total_power is the
// list's total selection weight.
std::vector<cn_light_record> build_light_table(const std::vector<EmissiveTriangle>& tris,
                                               const cn_delta_light* deltas, uint32_t delta_count,
                                               double* total_power);

struct EnvironmentTables {
    std::vector<float> marginal;     // height + 1
    std::vector<float> conditional;  // height * (width + 1)
    std::vector<float> pdfs;         // height * width
    double power = 0.0;
    // Search guides for the CDFs: marginal_buckets + 1 entries, then
    // conditional_buckets + 1 per row (see build_cdf_guide).
    std::vector<uint32_t> marginal_guide, conditional_guide;
    uint32_t marginal_buckets = 1, conditional_buckets = 1;
};

// environment.rs:130-218
EnvironmentTables build_environment_tables(uint32_t width, uint32_t height, const float* texels);

// The index the CDF inversion of environment.slang:41-58 returns for u is
// the last i in [0, count) with cdf[i] <= u (cdf[0] = 0); it never decreases
// as u grows. The guide holds that index at u = g / buckets for g in
// 0..buckets (buckets a power of two, so u * buckets and g / buckets are
// exact): for u in bucket g the answer lies in [guide[g], guide[g + 1]], and
// a binary search over that range returns the same index as one over the
// whole table. Returns false, writing nothing, for a table that does not
// rise monotonically.
uint32_t guide_buckets(uint32_t count);
bool build_cdf_guide(const float* cdf, uint32_t count, uint32_t buckets, uint32_t* guide);

// scene/mod.rs:665-671
float select_probability(double env_power, double light_power);

// glam's Mat4::transform_point3 over the rows of a 3x4 affine transform.
v3 transform_point(const float rows[12], v3 p);

// The inverse of a 3x4 affine transform, computed in f64 and rounded.
// Returns false when the transform is singular or not finite.
bool invert_affine(const float rows[12], float out[12]);

// Camera::basis (scene/mod.rs:448-461) for `aspect`, with the thin-lens
// pre-scaling of wavefront.rs:643-653 and the focus default of
// scene/lower.rs:227-232. Returns false on a degenerate camera.
bool camera_view(const cn_camera& camera, float aspect, CameraView* out);

}  // namespace cn
