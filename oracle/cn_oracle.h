/*
 * cn_oracle.h — TEST INFRASTRUCTURE ONLY. The CPU oracle of the wavefront
 * path tracer: a scalar restatement of the reference estimator
 * (the crates/cenote/shaders/ kernels) and of its host table builders
 * (crates/cenote/src/{lights,environment,color}.rs, scene/mod.rs).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker. The product never calls it.
 *
 * Scenes come in through the boundary's own data format (cn_scene_desc in
 * include/cenote_hip.h); nothing else of the product is used.
 */
#ifndef CN_ORACLE_H
#define CN_ORACLE_H

#include <stdint.h>

#include "../include/cenote_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* shaders/rng.slang:140-152 */
void orc_sample_2d(uint32_t pixel, uint32_t sample, uint32_t dimension, float out[2]);
float orc_sample_1d(uint32_t pixel, uint32_t sample, uint32_t dimension);
/* shaders/rng.slang:128-132 */
float orc_hash_uniform(uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t crossing);
/* Stochastic opacity's per-crossing trial and the shadow attenuation's
 * fixed-point log form (the order-free restatements of intersect.slang:64-76
 * and trace_shadow.slang:52-65). */
float orc_crossing_uniform(uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t instance, uint32_t primitive);
int32_t orc_log2_transmittance_fixed(float opacity);
float orc_exp2_fixed(int32_t sum);
/* The raw u32 pieces, for known-answer checks. */
uint32_t orc_lowbias32(uint32_t x);
uint32_t orc_laine_karras(uint32_t x, uint32_t seed);
uint32_t orc_sobol(uint32_t index, uint32_t axis);
uint32_t orc_reverse_bits(uint32_t x);

/* lights.rs:175-224 over explicit triangle and delta inputs:
 * tri_corners = 9 floats per triangle (world space), tri_emission = 3 per
 * triangle, tri_ids = (instance, primitive) pairs. Writes `n` records. */
void orc_build_lights(const float* tri_corners, const float* tri_emission,
                      const uint32_t* tri_ids, uint32_t tri_count,
                      const cn_delta_light* deltas, uint32_t delta_count,
                      cn_light_record* out, double* total_power);

/* environment.rs:130-218 over RGBA texels. marginal has h+1 floats,
 * conditional h*(w+1), pdfs h*w. */
void orc_env_tables(uint32_t width, uint32_t height, const float* texels, float* marginal,
                    float* conditional, float* pdfs, double* power);

/* scene/mod.rs:665-671 */
float orc_select_probability(double env_power, double light_power);

/* color.rs:29-31 and color.rs:13-17 */
float orc_luminance(const float rgb[3]);
void orc_acescg_from_rec709(const float rec709[3], float out[3]);

/* scene/mod.rs:448-461 with the lens scaling of wavefront.rs:643-653:
 * out = right[3], up[3], forward[3] */
void orc_camera_basis(const cn_camera* camera, float aspect, float out[9]);

orc_scene* orc_scene_create(const cn_scene_desc* desc, const float* bsdf_tables);
void orc_scene_destroy(orc_scene* scene);
uint32_t orc_scene_light_count(orc_scene* scene);
float orc_scene_select_prob(orc_scene* scene);
void orc_scene_light_records(orc_scene* scene, cn_light_record* out);

/* One wave (Wavefront::trace_then, wavefront.rs:607-629): the per-pixel
 * radiance accumulator and, when aov is nonzero, the AOV accumulators.
 * guide (RGBA per pixel) is the AOV feature-throughput scratch; it may be
 * NULL when aov is 0. Rows are split across `threads` host threads. */
void orc_trace(orc_scene* scene, uint32_t width, uint32_t height, uint32_t sample,
               uint32_t max_bounces, uint32_t light_sampling, int aov, float* radiance,
               float* albedo, float* normal, float* depth, int threads);

/* A sub-rectangle of one wave, for bounded CPU-baseline samples. */
void orc_trace_rows(orc_scene* scene, uint32_t width, uint32_t height, uint32_t row0,
                    uint32_t row1, uint32_t sample, uint32_t max_bounces,
                    uint32_t light_sampling, int aov, float* radiance, float* albedo,
                    float* normal, float* depth, int threads);

/* Film accumulation (shaders/accumulate.slang:46-67) of `count` waves
 * starting at sample index `first`, into sums (reset when first == 0). */
void orc_accumulate(orc_scene* scene, uint32_t width, uint32_t height, uint32_t first,
                    uint32_t count, uint32_t max_bounces, uint32_t light_sampling,
                    float* beauty_sum, float* albedo_sum, float* normal_sum, float* depth_sum,
                    int threads);

/* The beauty sums of samples first .. first+count-1 over rows [row0, row1)
 * of a width x height wave (reset when a pixel's first sample is 0);
 * threads split the band's pixels. */
void orc_accumulate_rows(orc_scene* scene, uint32_t width, uint32_t height, uint32_t row0, uint32_t row1,
                         uint32_t first, uint32_t count, uint32_t max_bounces, uint32_t light_sampling,
                         float* beauty_sum, int threads);

/* The beauty sums of samples first .. first+count-1 at each of `pixel_count`
 * listed pixel indices of a width x height wave, one 4-float sum per entry. */
void orc_accumulate_pixels(orc_scene* scene, uint32_t width, uint32_t height, const uint32_t* pixels,
                           uint32_t pixel_count, uint32_t first, uint32_t count, uint32_t max_bounces,
                           uint32_t light_sampling, float* beauty_sum, int threads);

/* The bounce-0 closest hit per pixel (shaders/intersect.slang:45-88). */
void orc_first_hits(orc_scene* scene, uint32_t width, uint32_t height, uint32_t sample,
                    uint32_t* instance, uint32_t* primitive, float* bary, int threads);

/* Environment::sample (environment.slang:90-104) for `count` uniform pairs:
 * 12 floats per pair — direction xyz, pdf, radiance rgb, the inverted
 * (v, u) position, the row and column as u32 bits, zero. */
void orc_env_sample(orc_scene* scene, const float* u, uint32_t count, float* out);

#ifdef __cplusplus
}
#endif

#endif /* CN_ORACLE_H */
