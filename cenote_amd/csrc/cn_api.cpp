// cn_api.cpp — the C ABI (include/cenote_hip.h) over the gfx950 stage
// kernels: context, scene residency, the wavefront engine and the film.

#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cenote_hip.h"
#include "cn_bvh.h"
#include "cn_launch.h"
#include "cn_prep.h"
#include "cn_records.h"

using namespace cn;

namespace {

thread_local std::string g_last_error;

int fail(int status, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}

#define CN_HIP(call)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (call);                                                                            \
        if (e_ != hipSuccess) {                                                                            \
            return fail(e_ == hipErrorOutOfMemory ? CN_E_ALLOC : CN_E_DEVICE_FAULT, "%s: %s (%s:%d)", #call, \
                        hipGetErrorString(e_), __FILE__, __LINE__);                                        \
        }                                                                                                  \
    } while (0)

// A device allocation that frees itself.
template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t count = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    void reset() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    hipError_t alloc(size_t n) {
        reset();
        count = n;
        if (n == 0) return hipSuccess;
        return hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T));
    }
    hipError_t upload(const T* src, size_t n) {
        hipError_t e = alloc(n);
        if (e != hipSuccess || n == 0) return e;
        return hipMemcpy(ptr, src, n * sizeof(T), hipMemcpyHostToDevice);
    }
};

}  // namespace

struct cn_context {
    int device = 0;
    uint32_t compute_units = 0;
    std::string name;
    DevBuf<float> tables;  // the BSDF tables, shared by every scene
    bool tables_ready = false;
};

struct cn_scene {
    cn_context* ctx = nullptr;
    cn_camera camera{};
    DevBuf<BvhNode4> nodes;
    DevBuf<BvhTriangle> tris;
    DevBuf<InstanceRecord> instances;
    DevBuf<cn_material> materials;
    DevBuf<cn_light_record> lights;
    DevBuf<float> positions, normals;
    DevBuf<uint32_t> indices;
    DevBuf<float4> env_texels;
    DevBuf<float> env_marginal, env_conditional, env_pdfs;
    std::vector<cn_light_record> host_lights;
    EnvironmentTables host_env;
    cn_scene_info info{};
    SceneView view{};
};

namespace {

// Per-kernel timing: a HIP event pair on the launch stream around every
// launch, read back and summed on demand.
struct EventProfiler final : Profiler {
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, size_t>> open;  // (kind, start event index)
    std::vector<std::tuple<int, size_t, size_t>> pairs;
    size_t used = 0;
    double total_ms[K_KINDS] = {};
    uint64_t launches[K_KINDS] = {};

    hipEvent_t next() {
        if (used == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(int kind) override {
        hipEvent_t e = next();
        if (!e) return;
        (void)hipEventRecord(e, stream);
        open.emplace_back(kind, used - 1);
    }
    void end(int kind) override {
        if (open.empty()) return;
        const auto [k, start] = open.back();
        open.pop_back();
        hipEvent_t e = next();
        if (!e) return;
        (void)hipEventRecord(e, stream);
        pairs.emplace_back(kind == k ? k : kind, start, used - 1);
    }
    // Fold the recorded pairs into the totals (the stream must be idle).
    void collect() {
        for (const auto& [kind, a, b] : pairs) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, pool[a], pool[b]) == hipSuccess) {
                total_ms[kind] += ms;
                launches[kind] += 1;
            }
        }
        pairs.clear();
        open.clear();
        used = 0;
    }
    ~EventProfiler() override {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace

struct cn_renderer {
    cn_context* ctx = nullptr;
    std::unique_ptr<EventProfiler> profiler;
    uint32_t capacity = 0;
    uint32_t max_bounces = 0;
    uint32_t light_sampling = 0;
    hipStream_t stream = nullptr;
    LaunchShape shape{};
    int counting = 0;
    // path pool
    DevBuf<float4> origin, direction, throughput;
    DevBuf<uint4> hit;
    DevBuf<uint32_t> state, ray_queue0, ray_queue1, hit_queue, miss_queue;
    DevBuf<float4> shadow_origin, shadow_dir, shadow_contrib;
    DevBuf<uint32_t> shadow_prim, spill;
    DevBuf<unsigned long long> traversal;
    DevBuf<uint32_t> counters;
    // scratch targets for cn_trace / cn_trace_aov
    DevBuf<float4> radiance, albedo, normal, guide;
    DevBuf<float> depth;
    size_t target_pixels = 0;
    uint32_t last_ranges = 0;
    PathPool pool{};
};

struct cn_film {
    cn_context* ctx = nullptr;
    uint32_t width = 0, height = 0;
    uint32_t samples = 0;
    DevBuf<float4> beauty, beauty_sum, albedo, albedo_sum, normal, normal_sum, guide;
    DevBuf<float> depth, depth_sum;
    DevBuf<float4> out_beauty, out_albedo, out_normal;
    DevBuf<float> out_depth;
};

namespace {

bool read_tables(std::vector<float>* out, std::string* error) {
    // The blobs ship beside the library: cenote_amd/data/{ggx_energy,ltc_sheen}.bin.
    Dl_info info;
    std::string dir;
    if (dladdr(reinterpret_cast<void*>(&read_tables), &info) && info.dli_fname) {
        dir = info.dli_fname;
        const size_t slash = dir.find_last_of('/');
        dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
    }
    const char* env = std::getenv("CENOTE_HIP_DATA");
    const std::string base = env ? std::string(env) : dir + "/data";
    out->resize(BSDF_TABLE_FLOATS);
    const char* names[2] = {"ggx_energy.bin", "ltc_sheen.bin"};
    const size_t counts[2] = {13856, 3072};
    size_t offset = 0;
    for (int i = 0; i < 2; ++i) {
        const std::string path = base + "/" + names[i];
        FILE* f = std::fopen(path.c_str(), "rb");
        if (!f) {
            *error = "cannot open BSDF table " + path;
            return false;
        }
        const size_t got = std::fread(out->data() + offset, 4, counts[i], f);
        std::fclose(f);
        if (got != counts[i]) {
            *error = "short BSDF table " + path;
            return false;
        }
        offset += counts[i];
    }
    return true;
}

bool material_ok(const cn_material& m) {
    return m.base_color_texture == CN_TEXTURE_NONE && m.specular_roughness_texture == CN_TEXTURE_NONE &&
           m.metalness_texture == CN_TEXTURE_NONE && m.emission_texture == CN_TEXTURE_NONE &&
           m.opacity_texture == CN_TEXTURE_NONE && m.normal_texture == CN_TEXTURE_NONE;
}

}  // namespace

extern "C" {

const char* cn_last_error(void) { return g_last_error.c_str(); }

int cn_context_create(int device, cn_context** out) {
    if (!out) return fail(CN_E_INVALID, "cn_context_create: null output");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        return fail(CN_E_NO_DEVICE, "no HIP device is visible");
    }
    if (device < 0 || device >= count) return fail(CN_E_NO_DEVICE, "device %d out of range (%d visible)", device, count);
    CN_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    CN_HIP(hipGetDeviceProperties(&prop, device));
    auto ctx = std::make_unique<cn_context>();
    ctx->device = device;
    ctx->compute_units = uint32_t(prop.multiProcessorCount);
    ctx->name = prop.gcnArchName;
    std::vector<float> tables;
    std::string error;
    if (!read_tables(&tables, &error)) return fail(CN_E_INVALID, "%s", error.c_str());
    CN_HIP(ctx->tables.upload(tables.data(), tables.size()));
    *out = ctx.release();
    return CN_OK;
}

int cn_context_destroy(cn_context* ctx) {
    if (!ctx) return fail(CN_E_INVALID, "cn_context_destroy: null context");
    (void)hipSetDevice(ctx->device);
    delete ctx;
    return CN_OK;
}

int cn_context_describe(cn_context* ctx, char* name, size_t name_capacity, uint32_t* compute_units) {
    if (!ctx) return fail(CN_E_INVALID, "cn_context_describe: null context");
    if (name && name_capacity) {
        std::snprintf(name, name_capacity, "%s", ctx->name.c_str());
    }
    if (compute_units) *compute_units = ctx->compute_units;
    return CN_OK;
}

int cn_scene_create(cn_context* ctx, const cn_scene_desc* desc, cn_scene** out) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !desc || !out) return fail(CN_E_INVALID, "cn_scene_create: null argument");
    if (desc->instance_count == 0) return fail(CN_E_SCENE, "a scene needs at least one instance");
    if (!desc->environment.texels || desc->environment.width == 0 || desc->environment.height == 0) {
        return fail(CN_E_SCENE, "a scene needs an environment (use a 1x1 black one for none)");
    }
    CN_HIP(hipSetDevice(ctx->device));
    auto scene = std::make_unique<cn_scene>();
    scene->ctx = ctx;
    scene->camera = desc->camera;

    // Pool the meshes.
    std::vector<uint32_t> vbase(desc->mesh_count), ibase(desc->mesh_count);
    size_t vtotal = 0, itotal = 0;
    for (uint32_t m = 0; m < desc->mesh_count; ++m) {
        const cn_mesh& mesh = desc->meshes[m];
        if (!mesh.positions || !mesh.normals || !mesh.indices) return fail(CN_E_SCENE, "mesh %u lacks arrays", m);
        for (size_t k = 0; k < 3 * size_t(mesh.triangle_count); ++k) {
            if (mesh.indices[k] >= mesh.vertex_count) return fail(CN_E_SCENE, "mesh %u index out of range", m);
        }
        vbase[m] = uint32_t(vtotal);
        ibase[m] = uint32_t(itotal);
        vtotal += mesh.vertex_count;
        itotal += mesh.triangle_count;
    }
    std::vector<float> positions(3 * vtotal), normals(3 * vtotal);
    std::vector<uint32_t> indices(3 * itotal);
    for (uint32_t m = 0; m < desc->mesh_count; ++m) {
        const cn_mesh& mesh = desc->meshes[m];
        std::memcpy(&positions[3 * size_t(vbase[m])], mesh.positions, 12 * size_t(mesh.vertex_count));
        std::memcpy(&normals[3 * size_t(vbase[m])], mesh.normals, 12 * size_t(mesh.vertex_count));
        std::memcpy(&indices[3 * size_t(ibase[m])], mesh.indices, 12 * size_t(mesh.triangle_count));
    }

    // Instances, materials, emissive triangles, world triangles.
    std::vector<InstanceRecord> records(desc->instance_count);
    std::vector<cn_material> materials(desc->instance_count);
    std::vector<EmissiveTriangle> emitters;
    std::vector<WorldTriangle> world;
    world.reserve(itotal);
    for (uint32_t i = 0; i < desc->instance_count; ++i) {
        const cn_instance& inst = desc->instances[i];
        if (inst.mesh >= desc->mesh_count) return fail(CN_E_SCENE, "instance %u names mesh %u", i, inst.mesh);
        if (!material_ok(inst.material)) {
            return fail(CN_E_SCENE, "instance %u: material textures are not supported on this path", i);
        }
        float inverse[12];
        if (!invert_affine(inst.object_to_world, inverse)) {
            return fail(CN_E_SCENE, "instance %u: transform must be invertible", i);
        }
        InstanceRecord& r = records[i];
        std::memset(&r, 0, sizeof(r));
        for (int row = 0; row < 3; ++row) {
            const float* a = inst.object_to_world + 4 * row;
            const float* b = inverse + 4 * row;
            r.o2w[row] = make_float4(a[0], a[1], a[2], a[3]);
            r.w2o[row] = make_float4(b[0], b[1], b[2], b[3]);
        }
        const cn_mesh& mesh = desc->meshes[inst.mesh];
        const uint32_t mask = inst.camera_visible ? MASK_ALL : (MASK_ALL & ~MASK_CAMERA);
        const bool opaque = inst.material.opacity >= 1.0f;
        r.vertex_base = vbase[inst.mesh];
        r.index_base = ibase[inst.mesh];
        r.flags = mask | (opaque ? 0x100u : 0u);
        r.light = LIGHT_NONE;
        materials[i] = inst.material;
        const v3 emission = V3(inst.material.emission[0], inst.material.emission[1], inst.material.emission[2]);
        // scene/mod.rs:268-281: a nonzero emission makes every triangle a
        // light, in primitive order; the record points at the first.
        const bool emissive = emission.x != 0.0f || emission.y != 0.0f || emission.z != 0.0f;
        if (emissive && mesh.triangle_count > 0) r.light = uint32_t(emitters.size());
        const uint32_t tri_flags = (mask << 8) | (opaque ? 0u : TRI_NON_OPAQUE);
        for (uint32_t p = 0; p < mesh.triangle_count; ++p) {
            v3 w[3];
            for (int k = 0; k < 3; ++k) {
                const uint32_t vi = mesh.indices[3 * size_t(p) + k];
                w[k] = transform_point(inst.object_to_world,
                                       V3(mesh.positions[3 * size_t(vi)], mesh.positions[3 * size_t(vi) + 1],
                                          mesh.positions[3 * size_t(vi) + 2]));
            }
            world.push_back(WorldTriangle{w[0], w[1], w[2], i, p, tri_flags});
            if (emissive) emitters.push_back(EmissiveTriangle{{w[0], w[1], w[2]}, emission, i, p});
        }
    }
    for (uint32_t d = 0; d < desc->delta_light_count; ++d) {
        const uint32_t kind = desc->delta_lights[d].kind;
        if (kind != CN_LIGHT_DISTANT && kind != CN_LIGHT_POINT) return fail(CN_E_SCENE, "delta light %u kind", d);
    }

    double light_power = 0.0;
    scene->host_lights = build_light_table(emitters, desc->delta_lights, desc->delta_light_count, &light_power);
    scene->host_env = build_environment_tables(desc->environment.width, desc->environment.height,
                                               desc->environment.texels);
    const float select = select_probability(scene->host_env.power, light_power);

    Bvh4 bvh;
    std::string error;
    if (!build_bvh4(world, STACK_MAX, &bvh, &error)) return fail(CN_E_SCENE, "%s", error.c_str());

    // Upload.
    CN_HIP(scene->nodes.upload(bvh.nodes.data(), bvh.nodes.size()));
    CN_HIP(scene->tris.upload(bvh.tris.data(), bvh.tris.size()));
    CN_HIP(scene->instances.upload(records.data(), records.size()));
    CN_HIP(scene->materials.upload(materials.data(), materials.size()));
    if (scene->host_lights.empty()) {
        cn_light_record zero{};
        CN_HIP(scene->lights.upload(&zero, 1));
    } else {
        CN_HIP(scene->lights.upload(scene->host_lights.data(), scene->host_lights.size()));
    }
    CN_HIP(scene->positions.upload(positions.data(), positions.size()));
    CN_HIP(scene->normals.upload(normals.data(), normals.size()));
    CN_HIP(scene->indices.upload(indices.data(), indices.size()));
    const size_t texel_count = size_t(desc->environment.width) * desc->environment.height;
    CN_HIP(scene->env_texels.upload(reinterpret_cast<const float4*>(desc->environment.texels), texel_count));
    CN_HIP(scene->env_marginal.upload(scene->host_env.marginal.data(), scene->host_env.marginal.size()));
    CN_HIP(scene->env_conditional.upload(scene->host_env.conditional.data(), scene->host_env.conditional.size()));
    CN_HIP(scene->env_pdfs.upload(scene->host_env.pdfs.data(), scene->host_env.pdfs.size()));

    SceneView& v = scene->view;
    v.nodes = scene->nodes.ptr;
    v.tris = scene->tris.ptr;
    v.instances = scene->instances.ptr;
    v.materials = scene->materials.ptr;
    v.lights = scene->lights.ptr;
    v.positions = scene->positions.ptr;
    v.normals = scene->normals.ptr;
    v.indices = scene->indices.ptr;
    v.tables = ctx->tables.ptr;
    v.env_texels = scene->env_texels.ptr;
    v.env_marginal = scene->env_marginal.ptr;
    v.env_conditional = scene->env_conditional.ptr;
    v.env_pdfs = scene->env_pdfs.ptr;
    v.light_count = uint32_t(scene->host_lights.size());
    v.env_width = desc->environment.width;
    v.env_height = desc->environment.height;
    v.env_select_prob = select;

    cn_scene_info& info = scene->info;
    info.instance_count = desc->instance_count;
    info.triangle_count = uint32_t(world.size());
    info.light_count = v.light_count;
    info.env_width = v.env_width;
    info.env_height = v.env_height;
    info.env_select_prob = select;
    info.env_power = scene->host_env.power;
    info.bvh_node_count = uint32_t(bvh.nodes.size());
    info.bvh_width = 4;
    info.bvh_node_bytes = sizeof(BvhNode4);
    info.bvh_triangle_bytes = sizeof(BvhTriangle);
    info.bvh_max_depth = bvh.max_depth;
    info.bvh_build_seconds = bvh.build_seconds;
    info.prep_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    *out = scene.release();
    return CN_OK;
}

int cn_scene_destroy(cn_scene* scene) {
    if (!scene) return fail(CN_E_INVALID, "cn_scene_destroy: null scene");
    (void)hipSetDevice(scene->ctx->device);
    delete scene;
    return CN_OK;
}

int cn_scene_set_camera(cn_scene* scene, const cn_camera* camera) {
    if (!scene || !camera) return fail(CN_E_INVALID, "cn_scene_set_camera: null argument");
    CameraView probe;
    if (!camera_view(*camera, 1.0f, &probe)) return fail(CN_E_INVALID, "degenerate camera");
    scene->camera = *camera;
    return CN_OK;
}

int cn_scene_info_get(cn_scene* scene, cn_scene_info* info) {
    if (!scene || !info) return fail(CN_E_INVALID, "cn_scene_info_get: null argument");
    *info = scene->info;
    return CN_OK;
}

int cn_scene_light_records(cn_scene* scene, cn_light_record* out, uint32_t capacity, uint32_t* count) {
    if (!scene) return fail(CN_E_INVALID, "cn_scene_light_records: null scene");
    const uint32_t n = uint32_t(scene->host_lights.size());
    if (count) *count = n;
    if (out) {
        if (capacity < n) return fail(CN_E_INVALID, "light record buffer holds %u of %u", capacity, n);
        std::memcpy(out, scene->host_lights.data(), n * sizeof(cn_light_record));
    }
    return CN_OK;
}

int cn_scene_env_tables(cn_scene* scene, float* marginal, float* conditional, float* pdfs) {
    if (!scene) return fail(CN_E_INVALID, "cn_scene_env_tables: null scene");
    const EnvironmentTables& t = scene->host_env;
    if (marginal) std::memcpy(marginal, t.marginal.data(), t.marginal.size() * 4);
    if (conditional) std::memcpy(conditional, t.conditional.data(), t.conditional.size() * 4);
    if (pdfs) std::memcpy(pdfs, t.pdfs.data(), t.pdfs.size() * 4);
    return CN_OK;
}

int cn_renderer_create(cn_context* ctx, uint32_t capacity, uint32_t max_bounces, uint32_t light_sampling,
                       cn_renderer** out) {
    if (!ctx || !out) return fail(CN_E_INVALID, "cn_renderer_create: null argument");
    if (capacity == 0) capacity = CN_DEFAULT_CAPACITY;
    if (max_bounces == 0 || max_bounces > 255) return fail(CN_E_INVALID, "max_bounces %u outside 1..255", max_bounces);
    if (light_sampling > CN_LIGHT_SAMPLING_NEE) return fail(CN_E_INVALID, "light_sampling %u", light_sampling);
    CN_HIP(hipSetDevice(ctx->device));
    auto r = std::make_unique<cn_renderer>();
    r->ctx = ctx;
    r->capacity = capacity;
    r->max_bounces = max_bounces;
    r->light_sampling = light_sampling;
    const uint32_t cus = std::max(1u, ctx->compute_units);
    r->shape.trace_blocks = cus * 8;
    r->shape.shade_blocks = cus * 2;
    r->shape.light_blocks = cus * 4;
    CN_HIP(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    CN_HIP(r->origin.alloc(capacity));
    CN_HIP(r->direction.alloc(capacity));
    CN_HIP(r->throughput.alloc(capacity));
    CN_HIP(r->hit.alloc(capacity));
    CN_HIP(r->state.alloc(capacity));
    CN_HIP(r->ray_queue0.alloc(capacity));
    CN_HIP(r->ray_queue1.alloc(capacity));
    CN_HIP(r->hit_queue.alloc(capacity));
    CN_HIP(r->miss_queue.alloc(capacity));
    CN_HIP(r->shadow_origin.alloc(capacity));
    CN_HIP(r->shadow_dir.alloc(capacity));
    CN_HIP(r->shadow_contrib.alloc(capacity));
    CN_HIP(r->shadow_prim.alloc(capacity));
    CN_HIP(r->spill.alloc(size_t(r->shape.trace_blocks) * 256 * STACK_SPILL));
    CN_HIP(r->traversal.alloc(4));
    CN_HIP(hipMemset(r->traversal.ptr, 0, 4 * sizeof(unsigned long long)));
    CN_HIP(r->counters.alloc(size_t(MAX_RANGES) * counters_per_range(max_bounces)));
    CN_HIP(hipMemset(r->counters.ptr, 0, r->counters.count * 4));
    PathPool& p = r->pool;
    p.origin = r->origin.ptr;
    p.direction = r->direction.ptr;
    p.hit = r->hit.ptr;
    p.throughput = r->throughput.ptr;
    p.state = r->state.ptr;
    p.ray_queue[0] = r->ray_queue0.ptr;
    p.ray_queue[1] = r->ray_queue1.ptr;
    p.hit_queue = r->hit_queue.ptr;
    p.miss_queue = r->miss_queue.ptr;
    p.shadow_origin = r->shadow_origin.ptr;
    p.shadow_dir = r->shadow_dir.ptr;
    p.shadow_contrib = r->shadow_contrib.ptr;
    p.shadow_prim = r->shadow_prim.ptr;
    p.stack_overflow = r->spill.ptr;
    p.traversal_counts = r->traversal.ptr;
    *out = r.release();
    return CN_OK;
}

int cn_renderer_destroy(cn_renderer* r) {
    if (!r) return fail(CN_E_INVALID, "cn_renderer_destroy: null renderer");
    (void)hipSetDevice(r->ctx->device);
    if (r->stream) {
        (void)hipStreamSynchronize(r->stream);
        (void)hipStreamDestroy(r->stream);
    }
    delete r;
    return CN_OK;
}

int cn_renderer_set_counters(cn_renderer* r, int enabled) {
    if (!r) return fail(CN_E_INVALID, "cn_renderer_set_counters: null renderer");
    r->counting = enabled ? 1 : 0;
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipMemsetAsync(r->traversal.ptr, 0, 4 * sizeof(unsigned long long), r->stream));
    CN_HIP(hipStreamSynchronize(r->stream));
    return CN_OK;
}

int cn_renderer_set_profiling(cn_renderer* r, int enabled) {
    if (!r) return fail(CN_E_INVALID, "cn_renderer_set_profiling: null renderer");
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    if (enabled) {
        r->profiler = std::make_unique<EventProfiler>();
        r->profiler->stream = r->stream;
    } else {
        r->profiler.reset();
    }
    return CN_OK;
}

int cn_renderer_kernel_times(cn_renderer* r, cn_kernel_times* out) {
    if (!r || !out) return fail(CN_E_INVALID, "cn_renderer_kernel_times: null argument");
    if (!r->profiler) return fail(CN_E_INVALID, "profiling is off");
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    r->profiler->collect();
    static const char* names[K_KINDS] = {"raygen", "intersect", "shade_miss", "shade_surface", "trace_shadow",
                                         "accumulate"};
    std::memset(out, 0, sizeof(*out));
    out->kernels = K_KINDS;
    for (int k = 0; k < K_KINDS; ++k) {
        std::snprintf(out->names[k], sizeof(out->names[k]), "%s", names[k]);
        out->total_ms[k] = r->profiler->total_ms[k];
        out->launches[k] = r->profiler->launches[k];
        r->profiler->total_ms[k] = 0.0;
        r->profiler->launches[k] = 0;
    }
    return CN_OK;
}

int cn_renderer_traversal_counts(cn_renderer* r, cn_traversal_counts* out) {
    if (!r || !out) return fail(CN_E_INVALID, "cn_renderer_traversal_counts: null argument");
    unsigned long long c[4];
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    CN_HIP(hipMemcpy(c, r->traversal.ptr, sizeof(c), hipMemcpyDeviceToHost));
    out->ray_nodes = c[0];
    out->ray_triangles = c[1];
    out->shadow_nodes = c[2];
    out->shadow_triangles = c[3];
    return CN_OK;
}

int cn_renderer_wave_counts(cn_renderer* r, cn_wave_counts* out) {
    if (!r || !out) return fail(CN_E_INVALID, "cn_renderer_wave_counts: null argument");
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    const uint32_t per = counters_per_range(r->max_bounces);
    std::vector<uint32_t> c(size_t(per) * MAX_RANGES);
    CN_HIP(hipMemcpy(c.data(), r->counters.ptr, c.size() * 4, hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    out->bounces = r->max_bounces;
    for (uint32_t range = 0; range < r->last_ranges; ++range) {
        const uint32_t* rc = &c[size_t(range) * per];
        for (uint32_t b = 0; b < r->max_bounces && b < 256; ++b) {
            out->rays[b] += rc[b * Q_COUNT + Q_RAY];
            out->hits[b] += rc[b * Q_COUNT + Q_HIT];
            out->misses[b] += rc[b * Q_COUNT + Q_MISS];
            out->shadows[b] += rc[b * Q_COUNT + Q_SHADOW];
        }
    }
    return CN_OK;
}

}  // extern "C"

namespace {

// One wave (Wavefront::trace_then, wavefront.rs:607-629): per pool-sized
// pixel range, raygen then the bounce rounds; all stream-ordered.
int enqueue_wave(cn_renderer* r, cn_scene* s, const WaveTargets& tg, uint32_t width, uint32_t height,
                 uint32_t sample) {
    CameraView cam;
    if (!camera_view(s->camera, float(width) / float(height), &cam)) return fail(CN_E_INVALID, "degenerate camera");
    const uint64_t pixels = uint64_t(width) * height;
    const uint64_t ranges = (pixels + r->capacity - 1) / r->capacity;
    if (ranges > MAX_RANGES) return fail(CN_E_INVALID, "target needs %llu pool ranges (max %u)", (unsigned long long)ranges, MAX_RANGES);
    const uint32_t per = counters_per_range(r->max_bounces);
    for (uint64_t range = 0; range < ranges; ++range) {
        const uint32_t base = uint32_t(range * r->capacity);
        const uint32_t count = uint32_t(std::min<uint64_t>(r->capacity, pixels - base));
        uint32_t* counters = r->counters.ptr + range * per;
        CN_HIP(launch_raygen(r->stream, cam, r->pool, tg, counters, per, width, height, base, count, sample, r->profiler.get()));
        for (uint32_t b = 0; b < r->max_bounces; ++b) {
            CN_HIP(launch_bounce(r->stream, r->shape, s->view, r->pool, tg, counters, base, count, b, r->max_bounces,
                                 r->light_sampling, sample, r->counting, r->profiler.get()));
        }
    }
    r->last_ranges = uint32_t(ranges);
    return CN_OK;
}

int ensure_targets(cn_renderer* r, size_t pixels, bool aov) {
    if (r->radiance.count < pixels) CN_HIP(r->radiance.alloc(pixels));
    if (aov) {
        if (r->albedo.count < pixels) CN_HIP(r->albedo.alloc(pixels));
        if (r->normal.count < pixels) CN_HIP(r->normal.alloc(pixels));
        if (r->guide.count < pixels) CN_HIP(r->guide.alloc(pixels));
        if (r->depth.count < pixels) CN_HIP(r->depth.alloc(pixels));
    }
    return CN_OK;
}

int check_target(uint32_t width, uint32_t height) {
    if (width == 0 || height == 0) return fail(CN_E_INVALID, "zero-sized trace target");
    return CN_OK;
}

}  // namespace

extern "C" {

int cn_trace_aov(cn_renderer* r, cn_scene* s, float* radiance, float* albedo, float* normal, float* depth,
                 uint32_t width, uint32_t height, uint32_t sample) {
    if (!r || !s) return fail(CN_E_INVALID, "cn_trace: null renderer or scene");
    if (int e = check_target(width, height)) return e;
    CN_HIP(hipSetDevice(r->ctx->device));
    const size_t pixels = size_t(width) * height;
    const bool aov = albedo || normal || depth;
    if (int e = ensure_targets(r, pixels, aov)) return e;
    WaveTargets tg{};
    tg.radiance = r->radiance.ptr;
    if (aov) {
        tg.albedo = r->albedo.ptr;
        tg.normal = r->normal.ptr;
        tg.depth = r->depth.ptr;
        tg.guide = r->guide.ptr;
    }
    if (int e = enqueue_wave(r, s, tg, width, height, sample)) return e;
    CN_HIP(hipStreamSynchronize(r->stream));
    if (radiance) CN_HIP(hipMemcpy(radiance, tg.radiance, pixels * 16, hipMemcpyDeviceToHost));
    if (albedo) CN_HIP(hipMemcpy(albedo, tg.albedo, pixels * 16, hipMemcpyDeviceToHost));
    if (normal) CN_HIP(hipMemcpy(normal, tg.normal, pixels * 16, hipMemcpyDeviceToHost));
    if (depth) CN_HIP(hipMemcpy(depth, tg.depth, pixels * 4, hipMemcpyDeviceToHost));
    return CN_OK;
}

int cn_trace(cn_renderer* r, cn_scene* s, float* radiance, uint32_t width, uint32_t height, uint32_t sample) {
    return cn_trace_aov(r, s, radiance, nullptr, nullptr, nullptr, width, height, sample);
}

int cn_first_hits(cn_renderer* r, cn_scene* s, uint32_t width, uint32_t height, uint32_t sample, uint32_t* instance,
                  uint32_t* primitive, float* bary) {
    if (!r || !s || !instance || !primitive || !bary) return fail(CN_E_INVALID, "cn_first_hits: null argument");
    if (int e = check_target(width, height)) return e;
    CN_HIP(hipSetDevice(r->ctx->device));
    CameraView cam;
    if (!camera_view(s->camera, float(width) / float(height), &cam)) return fail(CN_E_INVALID, "degenerate camera");
    const size_t pixels = size_t(width) * height;
    DevBuf<uint32_t> di, dp;
    DevBuf<float> db;
    CN_HIP(di.alloc(pixels));
    CN_HIP(dp.alloc(pixels));
    CN_HIP(db.alloc(2 * pixels));
    CN_HIP(launch_first_hits(r->stream, r->shape, s->view, cam, r->pool, width, height, sample, di.ptr, dp.ptr, db.ptr));
    CN_HIP(hipStreamSynchronize(r->stream));
    CN_HIP(hipMemcpy(instance, di.ptr, pixels * 4, hipMemcpyDeviceToHost));
    CN_HIP(hipMemcpy(primitive, dp.ptr, pixels * 4, hipMemcpyDeviceToHost));
    CN_HIP(hipMemcpy(bary, db.ptr, pixels * 8, hipMemcpyDeviceToHost));
    return CN_OK;
}

int cn_film_create(cn_context* ctx, uint32_t width, uint32_t height, cn_film** out) {
    if (!ctx || !out) return fail(CN_E_INVALID, "cn_film_create: null argument");
    if (width == 0 || height == 0) return fail(CN_E_INVALID, "zero-sized film");
    CN_HIP(hipSetDevice(ctx->device));
    auto f = std::make_unique<cn_film>();
    f->ctx = ctx;
    f->width = width;
    f->height = height;
    const size_t n = size_t(width) * height;
    CN_HIP(f->beauty.alloc(n));
    CN_HIP(f->beauty_sum.alloc(n));
    CN_HIP(f->albedo.alloc(n));
    CN_HIP(f->albedo_sum.alloc(n));
    CN_HIP(f->normal.alloc(n));
    CN_HIP(f->normal_sum.alloc(n));
    CN_HIP(f->guide.alloc(n));
    CN_HIP(f->depth.alloc(n));
    CN_HIP(f->depth_sum.alloc(n));
    *out = f.release();
    return CN_OK;
}

int cn_film_destroy(cn_film* f) {
    if (!f) return fail(CN_E_INVALID, "cn_film_destroy: null film");
    (void)hipSetDevice(f->ctx->device);
    delete f;
    return CN_OK;
}

int cn_film_reset(cn_film* f) {
    if (!f) return fail(CN_E_INVALID, "cn_film_reset: null film");
    f->samples = 0;
    return CN_OK;
}

int cn_film_samples(cn_film* f, uint32_t* samples) {
    if (!f || !samples) return fail(CN_E_INVALID, "cn_film_samples: null argument");
    *samples = f->samples;
    return CN_OK;
}

}  // extern "C"

namespace {

FilmView film_view(cn_film* f) {
    return FilmView{f->beauty.ptr, f->beauty_sum.ptr, f->albedo.ptr, f->albedo_sum.ptr,
                    f->normal.ptr, f->normal_sum.ptr, f->depth.ptr, f->depth_sum.ptr};
}

int enqueue_accumulate(cn_renderer* r, cn_scene* s, cn_film* f) {
    WaveTargets tg{f->beauty.ptr, f->albedo.ptr, f->normal.ptr, f->depth.ptr, f->guide.ptr};
    if (int e = enqueue_wave(r, s, tg, f->width, f->height, f->samples)) return e;
    CN_HIP(launch_accumulate(r->stream, film_view(f), f->width * f->height, f->samples == 0 ? 1u : 0u, r->profiler.get()));
    f->samples += 1;
    return CN_OK;
}

}  // namespace

extern "C" {

int cn_accumulate(cn_renderer* r, cn_scene* s, cn_film* f) {
    if (!r || !s || !f) return fail(CN_E_INVALID, "cn_accumulate: null argument");
    CN_HIP(hipSetDevice(r->ctx->device));
    if (int e = enqueue_accumulate(r, s, f)) return e;
    CN_HIP(hipStreamSynchronize(r->stream));
    return CN_OK;
}

int cn_accumulate_async(cn_renderer* r, cn_scene* s, cn_film* f, uint32_t count) {
    if (!r || !s || !f) return fail(CN_E_INVALID, "cn_accumulate_async: null argument");
    CN_HIP(hipSetDevice(r->ctx->device));
    for (uint32_t k = 0; k < count; ++k) {
        if (int e = enqueue_accumulate(r, s, f)) return e;
    }
    return CN_OK;
}

int cn_renderer_synchronize(cn_renderer* r) {
    if (!r) return fail(CN_E_INVALID, "cn_renderer_synchronize: null renderer");
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(hipStreamSynchronize(r->stream));
    return CN_OK;
}

int cn_film_sums(cn_film* f, float* beauty, float* albedo, float* normal, float* depth) {
    if (!f) return fail(CN_E_INVALID, "cn_film_sums: null film");
    CN_HIP(hipSetDevice(f->ctx->device));
    CN_HIP(hipDeviceSynchronize());
    const size_t n = size_t(f->width) * f->height;
    if (beauty) CN_HIP(hipMemcpy(beauty, f->beauty_sum.ptr, n * 16, hipMemcpyDeviceToHost));
    if (albedo) CN_HIP(hipMemcpy(albedo, f->albedo_sum.ptr, n * 16, hipMemcpyDeviceToHost));
    if (normal) CN_HIP(hipMemcpy(normal, f->normal_sum.ptr, n * 16, hipMemcpyDeviceToHost));
    if (depth) CN_HIP(hipMemcpy(depth, f->depth_sum.ptr, n * 4, hipMemcpyDeviceToHost));
    return CN_OK;
}

int cn_film_averages(cn_film* f, float* beauty, float* albedo, float* normal, float* depth) {
    if (!f) return fail(CN_E_INVALID, "cn_film_averages: null film");
    if (f->samples == 0) return fail(CN_E_INVALID, "averaging an empty film");
    const size_t n = size_t(f->width) * f->height;
    if (int e = cn_film_sums(f, beauty, albedo, normal, depth)) return e;
    // render/film.rs:194-197: sum / samples on the host.
    const float samples = float(f->samples);
    auto divide = [&](float* a, size_t count) {
        if (a) for (size_t i = 0; i < count; ++i) a[i] = a[i] / samples;
    };
    divide(beauty, 4 * n);
    divide(albedo, 4 * n);
    divide(normal, 4 * n);
    divide(depth, n);
    return CN_OK;
}

int cn_film_set_sample(cn_film* f, const float* beauty, const float* albedo, const float* normal,
                       const float* depth) {
    if (!f) return fail(CN_E_INVALID, "cn_film_set_sample: null film");
    CN_HIP(hipSetDevice(f->ctx->device));
    const size_t n = size_t(f->width) * f->height;
    if (beauty) CN_HIP(hipMemcpy(f->beauty.ptr, beauty, n * 16, hipMemcpyHostToDevice));
    if (albedo) CN_HIP(hipMemcpy(f->albedo.ptr, albedo, n * 16, hipMemcpyHostToDevice));
    if (normal) CN_HIP(hipMemcpy(f->normal.ptr, normal, n * 16, hipMemcpyHostToDevice));
    if (depth) CN_HIP(hipMemcpy(f->depth.ptr, depth, n * 4, hipMemcpyHostToDevice));
    return CN_OK;
}

int cn_film_accumulate_sample(cn_renderer* r, cn_film* f) {
    if (!r || !f) return fail(CN_E_INVALID, "cn_film_accumulate_sample: null argument");
    CN_HIP(hipSetDevice(r->ctx->device));
    CN_HIP(launch_accumulate(r->stream, film_view(f), f->width * f->height, f->samples == 0 ? 1u : 0u, nullptr));
    f->samples += 1;
    CN_HIP(hipStreamSynchronize(r->stream));
    return CN_OK;
}

int cn_resolve(cn_renderer* r, cn_film* f, float* beauty, float* albedo, float* normal, float* depth) {
    if (!r || !f) return fail(CN_E_INVALID, "cn_resolve: null argument");
    if (f->samples == 0) return fail(CN_E_INVALID, "resolving an empty film");
    CN_HIP(hipSetDevice(r->ctx->device));
    const size_t n = size_t(f->width) * f->height;
    if (!f->out_beauty.ptr) {
        CN_HIP(f->out_beauty.alloc(n));
        CN_HIP(f->out_albedo.alloc(n));
        CN_HIP(f->out_normal.alloc(n));
        CN_HIP(f->out_depth.alloc(n));
    }
    FilmView v{f->out_beauty.ptr, f->beauty_sum.ptr, f->out_albedo.ptr, f->albedo_sum.ptr,
               f->out_normal.ptr, f->normal_sum.ptr, f->out_depth.ptr, f->depth_sum.ptr};
    CN_HIP(launch_resolve(r->stream, v, uint32_t(n), float(f->samples)));
    CN_HIP(hipStreamSynchronize(r->stream));
    if (beauty) CN_HIP(hipMemcpy(beauty, f->out_beauty.ptr, n * 16, hipMemcpyDeviceToHost));
    if (albedo) CN_HIP(hipMemcpy(albedo, f->out_albedo.ptr, n * 16, hipMemcpyDeviceToHost));
    if (normal) CN_HIP(hipMemcpy(normal, f->out_normal.ptr, n * 16, hipMemcpyDeviceToHost));
    if (depth) CN_HIP(hipMemcpy(depth, f->out_depth.ptr, n * 4, hipMemcpyDeviceToHost));
    return CN_OK;
}

int cn_sample_sequence(cn_context* ctx, uint32_t pixel, uint32_t dimension, uint32_t count, float* points2d,
                       float* values1d) {
    if (!ctx || !points2d || !values1d) return fail(CN_E_INVALID, "cn_sample_sequence: null argument");
    if (count == 0) return CN_OK;
    CN_HIP(hipSetDevice(ctx->device));
    DevBuf<float> p, v;
    CN_HIP(p.alloc(2 * size_t(count)));
    CN_HIP(v.alloc(count));
    CN_HIP(launch_sample_sequence(nullptr, pixel, dimension, count, p.ptr, v.ptr));
    CN_HIP(hipDeviceSynchronize());
    CN_HIP(hipMemcpy(points2d, p.ptr, 8 * size_t(count), hipMemcpyDeviceToHost));
    CN_HIP(hipMemcpy(values1d, v.ptr, 4 * size_t(count), hipMemcpyDeviceToHost));
    return CN_OK;
}

}  // extern "C"
