// cn_scene.cpp — the C ABI's context and scene halves
// (include/cenote_hip.h): device bring-up, and scene residency — the
// reference's Scene::new (scene/mod.rs:254-338) from flattened records.

#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <atomic>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "cn_bvh.h"
#include "cn_handles.h"

using namespace cn;

namespace cn {

namespace {
thread_local std::string g_last_error;
}  // namespace

int fail(int status, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}

const char* last_error_text() { return g_last_error.c_str(); }

}  // namespace cn

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

const char* cn_last_error(void) { return cn::last_error_text(); }

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
    if (int e = static_arena_init(device)) return e;
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
    static std::atomic<uint64_t> serials{1};
    scene->serial = serials.fetch_add(1);
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
        r.pad[0] = material_class(inst.material);
        r.light = LIGHT_NONE;
        materials[i] = inst.material;
        const v3 emission = V3(inst.material.emission[0], inst.material.emission[1], inst.material.emission[2]);
        // scene/mod.rs:268-281: a nonzero emission makes every triangle a
        // light, in primitive order; the record points at the first.
        const bool emissive = emission.x != 0.0f || emission.y != 0.0f || emission.z != 0.0f;
        if (emissive && mesh.triangle_count > 0) r.light = uint32_t(emitters.size());
        const uint32_t tri_flags =
            (r.pad[0] << TRI_CLASS_SHIFT) | (mask << 8) | (opaque ? 0u : TRI_NON_OPAQUE);
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

    if (desc->bvh_builder != CN_BVH_DEVICE && desc->bvh_builder != CN_BVH_HOST)
        return fail(CN_E_INVALID, "unknown BVH builder %u", desc->bvh_builder);
    Bvh4 bvh;
    std::string error;
    if (desc->bvh_builder == CN_BVH_HOST) {
        if (!build_bvh4(world, STACK_MAX, &bvh, &error)) return fail(CN_E_SCENE, "%s", error.c_str());
        CN_HIP(scene->tris.upload(bvh.tris.data(), bvh.tris.size()));
    } else {
        BvhTriangle* tris = nullptr;
        if (!build_bvh4_device(world, STACK_MAX, &bvh, &tris, &error)) return fail(CN_E_SCENE, "%s", error.c_str());
        scene->tris.adopt(tris, world.size());
    }

    // Upload.
    {
        std::vector<BvhNode4Q> quantized(bvh.nodes.size());
        quantize_bvh4(bvh.nodes.data(), bvh.nodes.size(), quantized.data());
        CN_HIP(scene->nodes.upload(quantized.data(), quantized.size()));
    }
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
    CN_HIP(scene->env_marginal_guide.upload(scene->host_env.marginal_guide.data(),
                                            scene->host_env.marginal_guide.size()));
    CN_HIP(scene->env_conditional_guide.upload(scene->host_env.conditional_guide.data(),
                                               scene->host_env.conditional_guide.size()));

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
    v.env_marginal_guide = scene->env_marginal_guide.ptr;
    v.env_conditional_guide = scene->env_conditional_guide.ptr;
    v.env_marginal_buckets = scene->host_env.marginal_buckets;
    v.env_conditional_buckets = scene->host_env.conditional_buckets;
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
    scene->host_nodes = std::move(bvh.nodes);
    info.bvh_width = 4;
    info.bvh_node_bytes = sizeof(BvhNode4Q);
    info.bvh_triangle_bytes = sizeof(BvhTriangle);
    info.bvh_max_depth = bvh.max_depth;
    info.bvh_builder = desc->bvh_builder;
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

int cn_scene_bvh(cn_scene* scene, void* nodes, uint32_t node_capacity, uint32_t* node_count, void* triangles,
                 uint32_t triangle_capacity, uint32_t* triangle_count) {
    if (!scene) return fail(CN_E_INVALID, "cn_scene_bvh: null scene");
    CN_HIP(hipSetDevice(scene->ctx->device));
    if (node_count) *node_count = uint32_t(scene->host_nodes.size());
    if (triangle_count) *triangle_count = uint32_t(scene->tris.count);
    if (nodes) {
        if (node_capacity < scene->host_nodes.size()) return fail(CN_E_INVALID, "node buffer too small");
        std::memcpy(nodes, scene->host_nodes.data(), scene->host_nodes.size() * sizeof(BvhNode4));
    }
    if (triangles) {
        if (triangle_capacity < scene->tris.count) return fail(CN_E_INVALID, "triangle buffer too small");
        CN_HIP(hipMemcpy(triangles, scene->tris.ptr, scene->tris.count * sizeof(BvhTriangle), hipMemcpyDeviceToHost));
    }
    return CN_OK;
}

int cn_bvh_quantize(const void* nodes, uint32_t count, void* out) {
    if (count && (!nodes || !out)) return fail(CN_E_INVALID, "cn_bvh_quantize: null argument");
    quantize_bvh4(static_cast<const BvhNode4*>(nodes), count, static_cast<BvhNode4Q*>(out));
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

int cn_scene_sample_environment(cn_scene* scene, const float* u, uint32_t count, float* out) {
    if (!scene || (count && (!u || !out))) return fail(CN_E_INVALID, "cn_scene_sample_environment: null argument");
    if (count == 0) return CN_OK;
    CN_HIP(hipSetDevice(scene->ctx->device));
    DevBuf<float2> du;
    DevBuf<float> dout;
    CN_HIP(du.upload(reinterpret_cast<const float2*>(u), count));
    CN_HIP(dout.alloc(size_t(count) * 12));
    CN_HIP(launch_env_sample(nullptr, scene->view, du.ptr, count, dout.ptr));
    CN_HIP(hipDeviceSynchronize());
    CN_HIP(hipMemcpy(out, dout.ptr, size_t(count) * 12 * 4, hipMemcpyDeviceToHost));
    return CN_OK;
}

}  // extern "C"
