/*
 * cenote_hip.h — the C ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * A drop-in for the hot path of the `cenote` crate: scene residency
 * (`Scene::new` / `Scene::prep`), the wavefront engine (`Wavefront::new`,
 * `Wavefront::trace`), and the film (`Film`, `Renderer::accumulate`,
 * `Renderer::resolve`, `Film::averages`). Every entry point takes plain
 * pointers and sizes; no entry point exposes a device runtime type.
 *
 * Conventions kept from the reference (crates/cenote/src/lib.rs:37-38,
 * shaders/raygen.slang:7): right-handed, Y-up, the camera looks -Z; images
 * are row-major RGBA f32 with pixel (0, 0) top-left; the beauty alpha is 1
 * exactly once per pixel per sample; depth is the camera-plane z, +inf on
 * a miss. Colors are ACEScg.
 *
 * Errors: every call returns a cn_status. CN_OK is 0; on any other value
 * cn_last_error() returns a thread-local message. The reference panics on
 * programmer bugs (wavefront.rs:493-498, :618-623; gpu/submit.rs:265-269);
 * this ABI returns CN_E_INVALID instead.
 *
 * Ownership: the caller owns every host array and may free it as soon as
 * the call that received it returns (the library copies). The library owns
 * device memory. Destroy films, renderers and scenes before their context
 * (the order crates/cenote/src/gpu/mod.rs:224-239 keeps).
 *
 * Threading: one context per process; a renderer or film is used by one
 * thread at a time. Calls are synchronous unless the name ends in _async,
 * in which case they are ordered on the renderer's stream.
 */
#ifndef CENOTE_HIP_H
#define CENOTE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; the mapping of crates/cenote/src/error.rs:8-76. */
typedef enum cn_status {
    CN_OK = 0,
    CN_E_NO_DEVICE = 1,    /* Error::NoCapableGpu */
    CN_E_ALLOC = 2,        /* Error::Allocation */
    CN_E_SCENE = 3,        /* Error::Scene: a rejected scene description */
    CN_E_INVALID = 4,      /* a programmer bug the reference panics on */
    CN_E_DEVICE_FAULT = 5  /* Error::Vulkan: the device reported a failure */
} cn_status;

/* LightSampling (wavefront.rs:428-439, shaders/lights.slang:32-34). */
enum {
    CN_LIGHT_SAMPLING_MIS = 0,
    CN_LIGHT_SAMPLING_BSDF = 1,
    CN_LIGHT_SAMPLING_NEE = 2
};

/* Delta-light kinds (lights.rs:57-73). */
enum {
    CN_LIGHT_DISTANT = 1,
    CN_LIGHT_POINT = 2
};

/* "No texture" in a material's texture slots (material.rs:23). */
#define CN_TEXTURE_NONE 0xffffffffu

/* Wavefront::DEFAULT_CAPACITY and DEFAULT_MAX_BOUNCES (wavefront.rs:462-471). */
#define CN_DEFAULT_CAPACITY (1u << 20)
#define CN_DEFAULT_MAX_BOUNCES 8u
/* The most paths a capacity of 0 lets a wave hold here: 2^27 (228 bytes of
 * path state and slot targets each, at most 30.6 GB, allocated as views
 * need it), so that one wave carries 64 samples of a 1080p view and even
 * the deep-bounce launches fill the chip. */
#define CN_POOL_CAPACITY (1u << 27)
/* The largest pool a renderer may ask for. */
#define CN_POOL_CAPACITY_MAX (1u << 27)
/* A new renderer's tail bounce (cn_launch_shape.tail_bounce): after bounce
 * 6's trace one launch finishes every live path. */
#define CN_DEFAULT_TAIL_BOUNCE 6u

/*
 * One instance's lowered OpenPBR surface, 144 bytes, field for field the
 * GPU record of crates/cenote/src/material.rs:28-116 (colors in ACEScg,
 * weights already clamped by lowering, scene/lower.rs:316-399). Texture
 * slots must be CN_TEXTURE_NONE: material textures are outside this path.
 */
typedef struct cn_material {
    float base_color[3];
    float base_roughness;
    float emission[3];
    float metalness;
    float specular_weight;
    float specular_roughness;
    float specular_ior;
    float transmission_weight;
    float transmission_color[3];
    float transmission_depth;
    float coat_color[3];
    float coat_weight;
    float coat_roughness;
    float coat_ior;
    float coat_darkening;
    float fuzz_weight;
    float fuzz_color[3];
    float fuzz_roughness;
    float opacity;
    uint32_t thin_walled;
    uint32_t base_color_texture;
    uint32_t specular_roughness_texture;
    uint32_t metalness_texture;
    uint32_t emission_texture;
    uint32_t opacity_texture;
    uint32_t normal_texture;
} cn_material;

/*
 * A triangle mesh in object space (scene/mod.rs:58-76): packed xyz
 * positions, one unit shading normal per vertex, one uv per vertex (may be
 * NULL: zeros), and counter-clockwise index triples.
 */
typedef struct cn_mesh {
    const float* positions;
    const float* normals;
    const float* uvs;
    const uint32_t* indices;
    uint32_t vertex_count;
    uint32_t triangle_count;
} cn_mesh;

/*
 * One placed object (scene/mod.rs:78-87 `Object`, plus the
 * camera-visibility flag of scene/mod.rs:394-400): the mesh it places, the
 * rows of its 3x4 object-to-world transform (must be invertible), and its
 * material. The instance's position in the array is its custom index.
 */
typedef struct cn_instance {
    uint32_t mesh;
    uint32_t camera_visible;
    float object_to_world[12];
    cn_material material;
} cn_instance;

/* A delta light (lights.rs:57-73): distant `vector` = unit travel
 * direction and `color` = irradiance; point `vector` = position and
 * `color` = intensity. ACEScg. */
typedef struct cn_delta_light {
    uint32_t kind;
    float vector[3];
    float color[3];
} cn_delta_light;

/* An equirect environment, RGBA f32 ACEScg texels, row 0 = +Y
 * (environment.rs:33-79). */
typedef struct cn_environment {
    uint32_t width;
    uint32_t height;
    const float* texels;
} cn_environment;

/* The camera (scene/mod.rs:397-447, with the lens resolved the way
 * scene/lower.rs:219-234 does): aperture_radius 0 is a pinhole;
 * focus_distance <= 0 focuses at look_at. */
typedef struct cn_camera {
    float position[3];
    float look_at[3];
    float up[3];
    float vfov_degrees;
    float aperture_radius;
    float focus_distance;
} cn_camera;

typedef struct cn_scene_desc {
    const cn_mesh* meshes;
    uint32_t mesh_count;
    const cn_instance* instances;
    uint32_t instance_count;
    const cn_delta_light* delta_lights;
    uint32_t delta_light_count;
    cn_environment environment;
    cn_camera camera;
    uint32_t bvh_builder;  /* CN_BVH_DEVICE (0) or CN_BVH_HOST: both grow the same tree */
    uint32_t pad;
} cn_scene_desc;

/* Who builds the scene's BVH (no reference counterpart: the reference's
 * driver builds its acceleration structures, gpu/accel.rs:167-269). The
 * device builder is the default; the host builder is the same algorithm on
 * the CPU, kept as a cross-check. */
enum { CN_BVH_DEVICE = 0, CN_BVH_HOST = 1 };

/* The 80-byte light record as the kernels read it (lights.rs:81-115). */
typedef struct cn_light_record {
    float anchor[3];
    float pdf;
    float edge1[3];
    float alias_threshold;
    float edge2[3];
    uint32_t alias_index;
    float emission[3];
    uint32_t instance;
    uint32_t kind;
    uint32_t primitive;
    uint32_t pad[2];
} cn_light_record;

/* Scene facts for tests and measurement. */
typedef struct cn_scene_info {
    uint32_t instance_count;
    uint32_t triangle_count;   /* flattened world triangles */
    uint32_t light_count;      /* records in the alias table */
    uint32_t env_width;
    uint32_t env_height;
    float env_select_prob;     /* scene/mod.rs:665-671 */
    double env_power;          /* environment.rs:203-210 */
    uint32_t bvh_node_count;
    uint32_t bvh_width;        /* children per node */
    uint32_t bvh_node_bytes;
    uint32_t bvh_triangle_bytes;
    uint32_t bvh_max_depth;
    uint32_t bvh_builder;      /* the CN_BVH_* that built it */
    double bvh_build_seconds;
    double prep_seconds;       /* the last create or update, host and device */
    uint32_t bvh8_node_count;  /* the 8-wide tree over the same leaves (0: none, too deep) */
    uint32_t bvh8_max_depth;
} cn_scene_info;

/* Per-bounce queue counts of the most recent wave (the audit of
 * wavefront.rs:874-913): rays traced, hits, misses, shadow rays. */
typedef struct cn_wave_counts {
    uint32_t bounces;
    uint32_t rays[256];
    uint32_t hits[256];
    uint32_t misses[256];
    uint32_t shadows[256];
} cn_wave_counts;

/* BVH traversal work of the waves traced since counters were last turned on. */
typedef struct cn_traversal_counts {
    uint64_t ray_nodes;      /* nodes fetched by closest-hit rays */
    uint64_t ray_triangles;  /* triangles tested by closest-hit rays */
    uint64_t shadow_nodes;
    uint64_t shadow_triangles;
} cn_traversal_counts;

/* Device time per stage kernel, summed over the launches since the last
 * read (HIP events recorded on the renderer's stream around each launch). */
typedef struct cn_kernel_times {
    uint32_t kernels;
    char names[8][24];
    double total_ms[8];
    uint64_t launches[8];
} cn_kernel_times;

typedef struct cn_context cn_context;
typedef struct cn_scene cn_scene;
typedef struct cn_renderer cn_renderer;
typedef struct cn_film cn_film;

/* Thread-local message of the last failing call on this thread. */
const char* cn_last_error(void);

/* Context::new (crates/cenote/src/gpu/mod.rs:93). */
int cn_context_create(int device, cn_context** out);
int cn_context_destroy(cn_context* ctx);
/* Device name and compute-unit count, for reports. */
int cn_context_describe(cn_context* ctx, char* name, size_t name_capacity,
                        uint32_t* compute_units);

/* Scene::new (scene/mod.rs:254-338): copies the description, lowers the
 * light list (lights.rs:175-224) and environment tables
 * (environment.rs:130-218) on the host, builds the BVH that replaces the
 * TLAS/BLAS pair (gpu/accel.rs:82-163), and uploads everything. */
int cn_scene_create(cn_context* ctx, const cn_scene_desc* desc, cn_scene** out);
int cn_scene_destroy(cn_scene* scene);
/* Camera edits between frames (Scene::camera_mut, scene/mod.rs:391-394). */
int cn_scene_set_camera(cn_scene* scene, const cn_camera* camera);

/* What an incremental update rebuilds: the parts of the description that
 * changed since the scene was created or last updated (the Dirty set of
 * scene/changeset.rs:351-356 folded to this ABI's granularity). */
enum {
    CN_DIRTY_CAMERA = 1,       /* the camera */
    CN_DIRTY_MATERIALS = 2,    /* instance materials and camera visibility (emission included) */
    CN_DIRTY_LIGHTS = 4,       /* the delta lights */
    CN_DIRTY_ENVIRONMENT = 8,  /* the environment map */
    CN_DIRTY_GEOMETRY = 16,    /* meshes, instances or transforms: everything is rebuilt */
    CN_DIRTY_ALL = 31
};

/* Scene::update (scene/prep.rs:122-177): rebuilds exactly what `dirty`
 * names from `desc`, the whole edited description, and leaves the rest of
 * the residency in place. A material edit rewrites the instance and
 * material tables, the triangles' flags in place in the BVH and the light
 * list; a delta-light edit the light list; an environment edit its tables;
 * each refreshes the light-selection probability. Without
 * CN_DIRTY_GEOMETRY, `desc` must place the same meshes (by vertex and
 * triangle count) with the same instances and transforms as the resident
 * scene, or the call fails with CN_E_SCENE. CN_E_SCENE leaves the scene
 * untouched (everything fallible runs before the first change); any other
 * failure is a device fault. The update waits for the device to go idle
 * before it replaces buffers, so it belongs between waves; the scene then
 * renders exactly as a scene freshly created from `desc` would. */
int cn_scene_update(cn_scene* scene, const cn_scene_desc* desc, uint32_t dirty);
int cn_scene_info_get(cn_scene* scene, cn_scene_info* info);
/* The BVH as raw records: the full-precision 128-byte nodes (six float4 of
 * child slab bounds, child refs, meta) and the 48-byte leaf-order triangles
 * the traversal reads (either output may be NULL; *count gets the record
 * count). The traversal reads each node quantized to 64 bytes, as
 * cn_bvh_quantize computes it. */
int cn_scene_bvh(cn_scene* scene, void* nodes, uint32_t node_capacity, uint32_t* node_count, void* triangles,
                 uint32_t triangle_capacity, uint32_t* triangle_count);
/* The 8-wide traversal nodes (128 bytes each: a grid origin and three
 * power-of-two steps, eight children's byte planes, eight child refs,
 * depth), collapsed from the same binary tree as the 4-wide nodes over the
 * same leaf-order triangles. *count = 0 when the scene has no 8-wide tree.
 * No reference counterpart: the reference's driver owns its acceleration
 * structure layout (gpu/accel.rs:167-269). */
int cn_scene_bvh8(cn_scene* scene, void* nodes, uint32_t node_capacity, uint32_t* node_count);
/* The 64-byte traversal nodes for `count` 128-byte nodes (host memory;
 * no device is involved): per node, an origin, three power-of-two grid
 * steps and each child's planes as bytes rounded outward, so every
 * quantized child box contains its full-precision box. */
int cn_bvh_quantize(const void* nodes, uint32_t count, void* out);
/* Host copies of the bit-exact tables, for parity checks. */
int cn_scene_light_records(cn_scene* scene, cn_light_record* out, uint32_t capacity,
                           uint32_t* count);
int cn_scene_env_tables(cn_scene* scene, float* marginal, float* conditional,
                        float* pdfs);
/* Environment importance sampling on the device for `count` uniform pairs
 * (u: 2 floats each) — Environment::sample (environment.slang:90-104), the
 * draw shade_surface's next-event estimation makes. out: 12 floats per
 * pair: direction xyz, pdf, radiance rgb, the inverted (v, u) position, and
 * the chosen row and column as u32 bits (the last float is zero). */
int cn_scene_sample_environment(cn_scene* scene, const float* u, uint32_t count, float* out);

/* Wavefront::new / Renderer::with_max_bounces (wavefront.rs:486-540,
 * render/mod.rs:151): capacity paths in the pool (0 = CN_POOL_CAPACITY,
 * at most CN_POOL_CAPACITY_MAX), max_bounces in 1..255, light_sampling a
 * CN_LIGHT_SAMPLING_* value. */
int cn_renderer_create(cn_context* ctx, uint32_t capacity, uint32_t max_bounces,
                       uint32_t light_sampling, cn_renderer** out);
int cn_renderer_destroy(cn_renderer* renderer);
/* Size the path pool and the wave's slot targets for waves of up to
 * `paths` paths (clamped to the capacity), so that later waves allocate
 * nothing. The reference allocates its whole pool up front
 * (wavefront.rs:486-540); here the pool otherwise grows as waves need it. */
int cn_renderer_reserve(cn_renderer* renderer, uint32_t paths, int aov);
/* Launch shape of the stage kernels (no reference counterpart: the
 * reference's dispatch sizes come from its queue headers). The autotuner
 * measures candidates and sets the fastest; every shape renders bitwise the
 * same images. */
enum {
    CN_SHADE_1024_LDS = 0,   /* 1024-lane groups, BSDF tables staged in LDS */
    CN_SHADE_512_LDS = 1,
    CN_SHADE_256_LDS = 2,
    CN_SHADE_256_GLOBAL = 3, /* 256-lane groups, tables read through the caches */
    /* As above, each group first ordering its hits by material class: */
    CN_SHADE_512_LDS_SORTED = 4,
    CN_SHADE_256_LDS_SORTED = 5,
    CN_SHADE_256_GLOBAL_SORTED = 6
};
typedef struct cn_launch_shape {
    uint32_t trace_blocks;   /* 256-lane workgroups of the persistent trace kernel, 1..65536 */
    uint32_t shade_variant;  /* a CN_SHADE_* value */
    uint32_t shade_blocks;   /* most workgroups of shade_surface, 1..65536 */
    uint32_t light_blocks;   /* most 256-lane workgroups of shade_miss, 1..65536 */
    uint32_t bvh_width;      /* 4 or 8: which of a scene's trees the traversal walks (a scene
                              * without an 8-wide tree is walked 4-wide) */
    uint32_t tail_bounce;    /* 0: every bounce as wavefront stages; b in 1..255: after bounce b's
                              * trace, one launch carries every live path through its remaining
                              * bounces (the same films, bit for bit; waves with counters on keep
                              * the stages) */
} cn_launch_shape;
int cn_renderer_get_shape(cn_renderer* renderer, cn_launch_shape* out);
int cn_renderer_set_shape(cn_renderer* renderer, const cn_launch_shape* shape);
/* Replay each wave's launch sequence from a captured hipGraph when a wave
 * of the same shape (views, sizes, sample counts per job, scene groups)
 * was captured before; progressive frames of the same views then cost one
 * graph launch each. Off while profiling or counting. No reference
 * counterpart (the reference records a fresh command buffer per wave,
 * wavefront.rs:607-629). */
int cn_renderer_set_graphs(cn_renderer* renderer, int enabled);
/* Waves captured into new graphs and waves replayed from the cache. */
int cn_renderer_graph_stats(cn_renderer* renderer, uint64_t* captures, uint64_t* replays);
/* Count BVH node and triangle fetches in later waves (slower). */
int cn_renderer_set_counters(cn_renderer* renderer, int enabled);
int cn_renderer_wave_counts(cn_renderer* renderer, cn_wave_counts* out);
int cn_renderer_traversal_counts(cn_renderer* renderer, cn_traversal_counts* out);
/* The same counts per bounce: out[4 * b + k] for b < `bounces` (at most
 * 256), k = ray nodes, ray triangles, shadow nodes, shadow triangles; the
 * shadow counts of bounce b are those of bounce b - 1's connections, which
 * bounce b's trace launch walks. */
int cn_renderer_traversal_counts_by_bounce(cn_renderer* renderer, uint64_t* out, uint32_t bounces);
/* Time every stage launch with HIP events (measurement only; the
 * reference has no counterpart — it carries no GPU timestamps). */
int cn_renderer_set_profiling(cn_renderer* renderer, int enabled);
/* Per-kernel device time since the last call; resets the totals. */
int cn_renderer_kernel_times(cn_renderer* renderer, cn_kernel_times* out);
/* Measurement only (no reference counterpart): one launch of `blocks`
 * 256-lane workgroups of a kernel whose waves each issue 8 * iterations VALU
 * instructions of one kind in eight independent chains — f32 fused
 * multiply-add, u32 add or f32 max — the yardsticks the SQ instruction
 * counters of the stage kernels are read against. *ms is the launch's device
 * time, *valu_per_wave its instruction count per wave (the counters add a
 * few setup instructions). */
enum { CN_CALIBRATE_FMA_F32 = 0, CN_CALIBRATE_ADD_U32 = 1, CN_CALIBRATE_MAX_F32 = 2 };
int cn_calibrate_valu(cn_context* ctx, uint32_t kind, uint32_t iterations, uint32_t blocks, double* ms,
                      uint64_t* valu_per_wave);

/* Wavefront::trace (wavefront.rs:571-582): one sample of a width x height
 * target into `radiance` (host RGBA f32, width*height*4 floats). */
int cn_trace(cn_renderer* renderer, cn_scene* scene, float* radiance, uint32_t width,
             uint32_t height, uint32_t sample);
/* As cn_trace, also returning the wave's AOV accumulators (any may be
 * NULL): albedo and normal RGBA f32, depth f32. */
int cn_trace_aov(cn_renderer* renderer, cn_scene* scene, float* radiance, float* albedo,
                 float* normal, float* depth, uint32_t width, uint32_t height,
                 uint32_t sample);
/* The bounce-0 hits of one wave (the first-hit identity check): instance
 * and primitive per pixel (0xffffffff on a miss) and barycentrics. */
int cn_first_hits(cn_renderer* renderer, cn_scene* scene, uint32_t width, uint32_t height,
                  uint32_t sample, uint32_t* instance, uint32_t* primitive, float* bary);

/* Film (render/film.rs:66-210). */
int cn_film_create(cn_context* ctx, uint32_t width, uint32_t height, cn_film** out);
int cn_film_destroy(cn_film* film);
/* Film::reset: the next accumulate overwrites (sample index 0). */
int cn_film_reset(cn_film* film);
int cn_film_samples(cn_film* film, uint32_t* samples);
/* Renderer::accumulate (render/mod.rs:245-259): one wave at sample index
 * film.samples, folded into the film's sums; film.samples += 1. */
int cn_accumulate(cn_renderer* renderer, cn_scene* scene, cn_film* film);
/* `count` consecutive accumulates, stream-ordered, no host sync between
 * them; the sums are bitwise those of `count` cn_accumulate calls. */
int cn_accumulate_async(cn_renderer* renderer, cn_scene* scene, cn_film* film,
                        uint32_t count);
/* Largest film (width * height) cn_accumulate_async_static takes. */
#define CN_STATIC_MAX_PIXELS (1u << 18)
/* cn_accumulate_async without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc). The wave's path
 * pool, queues and slot targets live in one library-owned device array of
 * 53 MiB per device (taken when a context is created on the device, held
 * until the process ends), so waves hold at most CN_STATIC_MAX_PIXELS
 * paths and a film of more pixels fails with CN_E_INVALID. The sums are
 * bitwise those cn_accumulate_async gives. Calls on one device share the
 * array: calls on one renderer are ordered by its stream, but calls whose
 * waves overlap in time on one device (two renderers' streams at once)
 * race. The renderer, scene and film must already exist. */
int cn_accumulate_async_static(cn_renderer* renderer, cn_scene* scene, cn_film* film,
                               uint32_t count);
/* One view of a batch: `samples` more samples of `camera` on `scene`,
 * folded into `film` (whose size sets the view's resolution), with paths
 * capped at `max_bounces` (0 = the renderer's; at most the renderer's). */
typedef struct cn_view_job {
    cn_scene* scene;
    cn_film* film;
    cn_camera camera;
    uint32_t samples;
    uint32_t max_bounces;
} cn_view_job;
/* Many views in shared waves (no reference counterpart: the reference
 * renders one view per Renderer::accumulate, render/mod.rs:245-259).
 * Each wave packs as many (view, sample) pairs as the path pool holds, in
 * view order; every film's sums are bitwise those of sequential
 * cn_accumulate calls on its view. Films must be distinct. Blocks until
 * done. */
int cn_render_batch(cn_renderer* renderer, const cn_view_job* views, uint32_t count);
/* --- Serving (no reference counterpart; the nearest is the progressive
 * Session, render/session.rs:480-489, which renders one view) -----------
 *
 * A request loop that admits progressive render requests while their films
 * fit an HBM budget, keeps scenes resident, and interleaves the admitted
 * requests into shared waves at sample granularity. Waves run on three lanes
 * of the server's own (renderers with their own streams and path pools,
 * copying the given renderer's depth, light sampling and launch shape), one
 * wave in flight per lane. The first-frame lane takes, as soon as it is
 * free, one sample of every admitted request that has none yet; the two
 * bulk lanes take waves of one more sample for every other unfinished
 * request plus leftover room as more samples in admission order. A wave's
 * accumulate waits for the earlier waves that fold into its films, so a
 * request's film receives its samples in index order and its result is
 * bitwise the same alone as in any mix. Times are seconds on the server's
 * steady clock. */
typedef struct cn_server cn_server;

enum {
    CN_REQUEST_QUEUED = 0,   /* waiting for budget */
    CN_REQUEST_ACTIVE = 1,   /* admitted, samples in progress */
    CN_REQUEST_DONE = 2,     /* target reached, result ready */
    CN_REQUEST_RELEASED = 3  /* result taken, film returned to the server */
};

typedef struct cn_request_desc {
    cn_scene* scene;
    cn_camera camera;
    uint32_t width, height;
    uint32_t spp;            /* target samples (with rel_mse: the most samples) */
    uint32_t max_bounces;    /* 0 = the renderer's; at most the renderer's */
    float rel_mse;           /* > 0: stop once the estimated relMSE of the mean
                              * luminance (variance of the mean over mean^2 + 1e-2,
                              * averaged over pixels) is at most this; 0 = spp only */
    float pad;
    double arrival_seconds;  /* the request's arrival on the server clock; < 0 = now */
} cn_request_desc;

typedef struct cn_request_status {
    uint64_t id;
    uint32_t state;               /* a CN_REQUEST_* value */
    uint32_t samples;             /* samples whose waves have completed */
    double arrival_seconds;
    double admitted_seconds;      /* -1 until admitted */
    double first_frame_seconds;   /* -1 until the first sample's wave completed */
    double done_seconds;          /* -1 until done */
    double rel_mse_estimate;      /* -1 without a noise target or before 2 samples */
} cn_request_status;

typedef struct cn_server_stats {
    uint64_t submitted, completed, waves, paths;
    uint32_t queued, active;
    uint64_t hbm_used_bytes;      /* films held by requests */
} cn_server_stats;

/* wave_paths: most paths per first-frame wave (0 = 2^24; at most the
 * renderer's capacity; smaller waves return first frames sooner); bulk waves
 * hold up to 4 x wave_paths while requests wait for their first frame and
 * 8 x wave_paths otherwise (at most the renderer's capacity). Each lane
 * allocates a pool of its largest waves' size (228 bytes per path) at
 * creation. */
int cn_server_create(cn_renderer* renderer, uint64_t hbm_budget_bytes, uint32_t wave_paths,
                     cn_server** out);
int cn_server_destroy(cn_server* server);
int cn_server_clock(cn_server* server, double* seconds);
int cn_server_submit(cn_server* server, const cn_request_desc* request, uint64_t* id);
/* One scheduling step: retire the waves that have completed, admit, and
 * enqueue a wave on every free lane that has work; when there was nothing to
 * do, wait up to a millisecond for a wave to complete. `busy` = queued +
 * active requests + waves in flight (0: idle). */
int cn_server_step(cn_server* server, uint32_t* busy);
/* Step until idle. Fails with CN_E_ALLOC, instead of waiting forever, when
 * queued requests cannot be admitted because finished requests' films
 * (results not yet taken) hold the budget. */
int cn_server_drain(cn_server* server);
int cn_server_status(cn_server* server, uint64_t id, cn_request_status* out);
int cn_server_stats_get(cn_server* server, cn_server_stats* out);
/* The averages (sums / samples) of a done request; releases its film. Any
 * output may be NULL. */
int cn_server_result(cn_server* server, uint64_t id, float* beauty, float* albedo, float* normal,
                     float* depth);

/* --- EXR output and input (output.rs:37-193) --------------------------
 * Linear ACEScg files with the AP1 chromaticities attribute; single-part
 * scanline, ZIP-compressed. cn_write_exr: row-major RGBA f32 (pixel (0,0)
 * top-left) as f32 R/G/B/A (write_exr, output.rs:37-57). cn_write_aov_exr:
 * the film's averages as one multi-layer file — f16 R/G/B/A, f32 Z, f16
 * albedo.R/G/B and normal.X/Y/Z (write_aov_exr, output.rs:84-136).
 * cn_read_exr: the first layer's R/G/B/A as f32, missing alpha = 1
 * (read_exr, output.rs:148-193); scanline or single-level tiled files,
 * NONE/RLE/ZIPS/ZIP, HALF/FLOAT/UINT channels. With rgba NULL it only
 * reports the size; capacity is in floats. */
int cn_write_exr(const char* path, uint32_t width, uint32_t height, const float* rgba);
int cn_write_aov_exr(const char* path, uint32_t width, uint32_t height, const float* beauty,
                     const float* albedo, const float* normal, const float* depth);
int cn_read_exr(const char* path, uint32_t* width, uint32_t* height, float* rgba, size_t capacity);

/* Block until the renderer's stream has drained. */
int cn_renderer_synchronize(cn_renderer* renderer);
/* Film::averages (render/film.rs:183-197): sums / samples on the host.
 * Any output may be NULL. */
int cn_film_averages(cn_film* film, float* beauty, float* albedo, float* normal,
                     float* depth);
/* The raw f32 sums (render/tests.rs reads these). Any output may be NULL. */
int cn_film_sums(cn_film* film, float* beauty, float* albedo, float* normal, float* depth);
/* Replace the film's sums (tests of the non-finite guard, render/tests.rs:1465-1520). */
int cn_film_set_sample(cn_film* film, const float* beauty, const float* albedo,
                       const float* normal, const float* depth);
/* Fold the film's current sample buffers into its sums (the accumulate
 * kernel alone, shaders/accumulate.slang:46-67); film.samples += 1. */
int cn_film_accumulate_sample(cn_renderer* renderer, cn_film* film);
/* Renderer::resolve (render/mod.rs:278-299): sums / samples on the device,
 * copied to the host outputs (any may be NULL). */
int cn_resolve(cn_renderer* renderer, cn_film* film, float* beauty, float* albedo,
               float* normal, float* depth);

/* The rng_test kernel (shaders/rng_test.slang:21-29): the first `count`
 * sample_2d points and sample_1d values of key (pixel, dimension),
 * computed on the device. points2d has 2*count floats. */
int cn_sample_sequence(cn_context* ctx, uint32_t pixel, uint32_t dimension, uint32_t count,
                       float* points2d, float* values1d);

#ifdef __cplusplus
}
#endif

#endif /* CENOTE_HIP_H */
