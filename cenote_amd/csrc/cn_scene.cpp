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

// The description's meshes pooled into one vertex and one index array: the
// first vertex and first triangle of each mesh in the pool.
struct MeshPool {
    std::vector<uint32_t> vbase, ibase;
    size_t vertices = 0, triangles = 0;
};

int pool_meshes(const cn_scene_desc* desc, MeshPool* pool) {
    pool->vbase.resize(desc->mesh_count);
    pool->ibase.resize(desc->mesh_count);
    for (uint32_t m = 0; m < desc->mesh_count; ++m) {
        const cn_mesh& mesh = desc->meshes[m];
        if (!mesh.positions || !mesh.normals || !mesh.indices) return fail(CN_E_SCENE, "mesh %u lacks arrays", m);
        for (size_t k = 0; k < 3 * size_t(mesh.triangle_count); ++k) {
            if (mesh.indices[k] >= mesh.vertex_count) return fail(CN_E_SCENE, "mesh %u index out of range", m);
        }
        pool->vbase[m] = uint32_t(pool->vertices);
        pool->ibase[m] = uint32_t(pool->triangles);
        pool->vertices += mesh.vertex_count;
        pool->triangles += mesh.triangle_count;
    }
    return CN_OK;
}

// The per-instance host phase (scene/mod.rs:259-281): in instance order,
// each instance's record (the GeometryRecord of scene/mod.rs:104-123),
// material, the flags its triangles carry in the BVH, and its emissive
// triangles in world space (a nonzero emission makes every triangle a light,
// in primitive order; the record points at the first). With `world`, every
// triangle in world space as well. Validates before anything is returned.
int prep_instances(const cn_scene_desc* desc, const MeshPool& pool, std::vector<InstanceRecord>* records,
                   std::vector<cn_material>* materials, std::vector<uint32_t>* tri_flags,
                   std::vector<EmissiveTriangle>* emitters, std::vector<WorldTriangle>* world) {
    records->assign(desc->instance_count, InstanceRecord{});
    materials->assign(desc->instance_count, cn_material{});
    tri_flags->assign(desc->instance_count, 0u);
    emitters->clear();
    if (world) world->reserve(pool.triangles);
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
        InstanceRecord& r = (*records)[i];
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
        r.vertex_base = pool.vbase[inst.mesh];
        r.index_base = pool.ibase[inst.mesh];
        r.flags = mask | (opaque ? 0x100u : 0u);
        r.material_class = material_class(inst.material);
        r.log2_transmittance = opaque ? 0 : log2_transmittance_fixed(inst.material.opacity);
        r.light = LIGHT_NONE;
        (*materials)[i] = inst.material;
        (*tri_flags)[i] = (r.material_class << TRI_CLASS_SHIFT) | (mask << 8) | (opaque ? 0u : TRI_NON_OPAQUE);
        const v3 emission = V3(inst.material.emission[0], inst.material.emission[1], inst.material.emission[2]);
        const bool emissive = emission.x != 0.0f || emission.y != 0.0f || emission.z != 0.0f;
        if (emissive && mesh.triangle_count > 0) r.light = uint32_t(emitters->size());
        if (!emissive && !world) continue;
        for (uint32_t p = 0; p < mesh.triangle_count; ++p) {
            v3 w[3];
            for (int k = 0; k < 3; ++k) {
                const uint32_t vi = mesh.indices[3 * size_t(p) + k];
                w[k] = transform_point(inst.object_to_world,
                                       V3(mesh.positions[3 * size_t(vi)], mesh.positions[3 * size_t(vi) + 1],
                                          mesh.positions[3 * size_t(vi) + 2]));
            }
            if (world) world->push_back(WorldTriangle{w[0], w[1], w[2], i, p, (*tri_flags)[i]});
            if (emissive) emitters->push_back(EmissiveTriangle{{w[0], w[1], w[2]}, emission, i, p});
        }
    }
    return CN_OK;
}

int check_delta_lights(const cn_scene_desc* desc) {
    if (desc->delta_light_count && !desc->delta_lights) return fail(CN_E_SCENE, "delta lights lack an array");
    for (uint32_t d = 0; d < desc->delta_light_count; ++d) {
        const uint32_t kind = desc->delta_lights[d].kind;
        if (kind != CN_LIGHT_DISTANT && kind != CN_LIGHT_POINT) return fail(CN_E_SCENE, "delta light %u kind", d);
    }
    return CN_OK;
}

int check_environment(const cn_environment& env) {
    if (!env.texels || env.width == 0 || env.height == 0) {
        return fail(CN_E_SCENE, "a scene needs an environment (use a 1x1 black one for none)");
    }
    return CN_OK;
}

// The light list (lights.rs:175-224) from the scene's emissive triangles and
// delta lights, kept on the host and uploaded.
int install_lights(cn_scene* scene, std::vector<EmissiveTriangle> emitters, const cn_delta_light* deltas,
                   uint32_t delta_count) {
    scene->host_emitters = std::move(emitters);
    scene->host_deltas.assign(deltas, deltas + delta_count);
    scene->host_lights = build_light_table(scene->host_emitters, scene->host_deltas.data(), delta_count,
                                           &scene->light_power);
    if (scene->host_lights.empty()) {
        cn_light_record zero{};
        CN_HIP(scene->lights.upload(&zero, 1));
    } else {
        CN_HIP(scene->lights.upload(scene->host_lights.data(), scene->host_lights.size()));
    }
    return CN_OK;
}

// The environment's texels and sampling tables (environment.rs:130-218).
int install_environment(cn_scene* scene, const cn_environment& env) {
    scene->host_env = build_environment_tables(env.width, env.height, env.texels);
    scene->env_width = env.width;
    scene->env_height = env.height;
    CN_HIP(scene->env_texels.upload(reinterpret_cast<const float4*>(env.texels), size_t(env.width) * env.height));
    const EnvironmentTables& t = scene->host_env;
    CN_HIP(scene->env_marginal.upload(t.marginal.data(), t.marginal.size()));
    CN_HIP(scene->env_conditional.upload(t.conditional.data(), t.conditional.size()));
    CN_HIP(scene->env_pdfs.upload(t.pdfs.data(), t.pdfs.size()));
    CN_HIP(scene->env_marginal_guide.upload(t.marginal_guide.data(), t.marginal_guide.size()));
    CN_HIP(scene->env_conditional_guide.upload(t.conditional_guide.data(), t.conditional_guide.size()));
    return CN_OK;
}

// The kernels' view of the scene and the parts of its info that follow the
// lights and the environment; run after every install.
void refresh_view(cn_scene* scene) {
    const float select = select_probability(scene->host_env.power, scene->light_power);
    SceneView& v = scene->view;
    v.nodes = scene->nodes.ptr;
    v.nodes8 = scene->nodes8.ptr;  // null without an 8-wide tree
    v.pnodes = scene->pnodes.ptr;
    v.tris = scene->tris.ptr;
    v.instances = scene->instances.ptr;
    v.materials = scene->materials.ptr;
    v.lights = scene->lights.ptr;
    v.positions = scene->positions.ptr;
    v.normals = scene->normals.ptr;
    v.indices = scene->indices.ptr;
    v.tables = scene->ctx->tables.ptr;
    v.env_texels = scene->env_texels.ptr;
    v.env_marginal = scene->env_marginal.ptr;
    v.env_conditional = scene->env_conditional.ptr;
    v.env_pdfs = scene->env_pdfs.ptr;
    v.env_marginal_guide = scene->env_marginal_guide.ptr;
    v.env_conditional_guide = scene->env_conditional_guide.ptr;
    v.env_marginal_buckets = scene->host_env.marginal_buckets;
    v.env_conditional_buckets = scene->host_env.conditional_buckets;
    v.light_count = uint32_t(scene->host_lights.size());
    v.env_width = scene->env_width;
    v.env_height = scene->env_height;
    v.env_select_prob = select;
    cn_scene_info& info = scene->info;
    info.light_count = v.light_count;
    info.env_width = v.env_width;
    info.env_height = v.env_height;
    info.env_select_prob = select;
    info.env_power = scene->host_env.power;
}

uint64_t next_serial() {
    static std::atomic<uint64_t> serials{1};
    return serials.fetch_add(1);
}

// Scene::new (scene/mod.rs:254-338) into an empty scene: validates the
// whole description before it uploads anything.
int prep_scene(cn_context* ctx, const cn_scene_desc* desc, cn_scene* scene) {
    if (desc->instance_count == 0) return fail(CN_E_SCENE, "a scene needs at least one instance");
    if (int e = check_environment(desc->environment)) return e;
    if (int e = check_delta_lights(desc)) return e;
    if (desc->bvh_builder != CN_BVH_DEVICE && desc->bvh_builder != CN_BVH_HOST)
        return fail(CN_E_INVALID, "unknown BVH builder %u", desc->bvh_builder);
    MeshPool pool;
    if (int e = pool_meshes(desc, &pool)) return e;
    std::vector<InstanceRecord> records;
    std::vector<cn_material> materials;
    std::vector<uint32_t> tri_flags;
    std::vector<EmissiveTriangle> emitters;
    std::vector<WorldTriangle> world;
    if (int e = prep_instances(desc, pool, &records, &materials, &tri_flags, &emitters, &world)) return e;

    scene->ctx = ctx;
    scene->camera = desc->camera;
    std::vector<float> positions(3 * pool.vertices), normals(3 * pool.vertices);
    std::vector<uint32_t> indices(3 * pool.triangles);
    for (uint32_t m = 0; m < desc->mesh_count; ++m) {
        const cn_mesh& mesh = desc->meshes[m];
        std::memcpy(&positions[3 * size_t(pool.vbase[m])], mesh.positions, 12 * size_t(mesh.vertex_count));
        std::memcpy(&normals[3 * size_t(pool.vbase[m])], mesh.normals, 12 * size_t(mesh.vertex_count));
        std::memcpy(&indices[3 * size_t(pool.ibase[m])], mesh.indices, 12 * size_t(mesh.triangle_count));
    }

    Bvh4 bvh;
    std::string error;
    if (desc->bvh_builder == CN_BVH_HOST) {
        if (!build_bvh4(world, STACK_MAX, &bvh, &error)) return fail(CN_E_SCENE, "%s", error.c_str());
        CN_HIP(scene->tris.upload(bvh.tris.data(), bvh.tris.size()));
        std::vector<BvhNode4Q> quantized(bvh.nodes.size());
        quantize_bvh4(bvh.nodes.data(), bvh.nodes.size(), quantized.data());
        CN_HIP(scene->nodes.upload(quantized.data(), quantized.size()));
        if (!bvh.nodes8.empty()) CN_HIP(scene->nodes8.upload(bvh.nodes8.data(), bvh.nodes8.size()));
    } else {
        // Built, collapsed and quantized on the device; the host keeps
        // copies only for read-back (cn_scene_bvh, cn_scene_bvh8).
        DeviceBvh dev;
        if (!build_bvh_device(world, STACK_MAX, &bvh, &dev, &error)) return fail(CN_E_SCENE, "%s", error.c_str());
        scene->tris.adopt(dev.tris, world.size());
        scene->nodes.adopt(dev.nodes, bvh.nodes.size());
        if (dev.nodes8) scene->nodes8.adopt(dev.nodes8, bvh.nodes8.size());
    }
    CN_HIP(scene->pnodes.alloc(scene->nodes.count));
    if (!decode_packet_nodes(scene->nodes.ptr, uint32_t(scene->nodes.count), scene->pnodes.ptr, &error))
        return fail(CN_E_SCENE, "%s", error.c_str());
    CN_HIP(scene->instances.upload(records.data(), records.size()));
    CN_HIP(scene->materials.upload(materials.data(), materials.size()));
    CN_HIP(scene->positions.upload(positions.data(), positions.size()));
    CN_HIP(scene->normals.upload(normals.data(), normals.size()));
    CN_HIP(scene->indices.upload(indices.data(), indices.size()));
    if (int e = install_lights(scene, std::move(emitters), desc->delta_lights, desc->delta_light_count)) return e;
    if (int e = install_environment(scene, desc->environment)) return e;
    scene->host_records = std::move(records);
    scene->host_instance_mesh.resize(desc->instance_count);
    for (uint32_t i = 0; i < desc->instance_count; ++i) scene->host_instance_mesh[i] = desc->instances[i].mesh;
    scene->host_mesh_shape.resize(2 * size_t(desc->mesh_count));
    for (uint32_t m = 0; m < desc->mesh_count; ++m) {
        scene->host_mesh_shape[2 * m] = desc->meshes[m].vertex_count;
        scene->host_mesh_shape[2 * m + 1] = desc->meshes[m].triangle_count;
    }
    refresh_view(scene);

    cn_scene_info& info = scene->info;
    info.instance_count = desc->instance_count;
    info.triangle_count = uint32_t(world.size());
    info.bvh_node_count = uint32_t(bvh.nodes.size());
    scene->host_nodes = std::move(bvh.nodes);
    info.bvh_width = 4;
    info.bvh_node_bytes = sizeof(BvhNode4Q);
    info.bvh_triangle_bytes = sizeof(BvhTriangle);
    info.bvh_max_depth = bvh.max_depth;
    info.bvh_builder = desc->bvh_builder;
    info.bvh_build_seconds = bvh.build_seconds;
    info.bvh8_node_count = uint32_t(bvh.nodes8.size());
    info.bvh8_max_depth = bvh.nodes8.empty() ? 0u : bvh.max_depth8;
    scene->host_nodes8 = std::move(bvh.nodes8);
    return CN_OK;
}

// True when the description places the same meshes (by vertex and triangle
// count) with the same instances and transforms as the resident scene.
bool same_geometry(const cn_scene* scene, const cn_scene_desc* desc) {
    if (desc->instance_count != scene->host_records.size()) return false;
    if (2 * size_t(desc->mesh_count) != scene->host_mesh_shape.size()) return false;
    for (uint32_t m = 0; m < desc->mesh_count; ++m) {
        if (desc->meshes[m].vertex_count != scene->host_mesh_shape[2 * m] ||
            desc->meshes[m].triangle_count != scene->host_mesh_shape[2 * m + 1])
            return false;
    }
    for (uint32_t i = 0; i < desc->instance_count; ++i) {
        const cn_instance& inst = desc->instances[i];
        if (inst.mesh != scene->host_instance_mesh[i]) return false;
        const InstanceRecord& r = scene->host_records[i];
        for (int row = 0; row < 3; ++row) {
            const float* a = inst.object_to_world + 4 * row;
            const float4 b = r.o2w[row];
            if (std::memcmp(a, &b, 16) != 0) return false;
        }
    }
    return true;
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
    CN_HIP(hipSetDevice(ctx->device));
    auto scene = std::make_unique<cn_scene>();
    if (int e = prep_scene(ctx, desc, scene.get())) return e;
    scene->serial = next_serial();
    scene->info.prep_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    *out = scene.release();
    return CN_OK;
}

int cn_scene_update(cn_scene* scene, const cn_scene_desc* desc, uint32_t dirty) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!scene || !desc) return fail(CN_E_INVALID, "cn_scene_update: null argument");
    if (dirty & ~uint32_t(CN_DIRTY_ALL)) return fail(CN_E_INVALID, "cn_scene_update: unknown dirty bits 0x%x", dirty);
    CN_HIP(hipSetDevice(scene->ctx->device));
    if (dirty & CN_DIRTY_CAMERA) {
        CameraView probe;
        if (!camera_view(desc->camera, 1.0f, &probe)) return fail(CN_E_SCENE, "degenerate camera");
    }
    if (dirty & CN_DIRTY_GEOMETRY) {
        // Meshes or placements changed: the BVH, the pooled geometry and
        // every table are rebuilt into a fresh scene, which replaces the
        // resident one only once it is whole.
        cn_scene fresh;
        if (int e = prep_scene(scene->ctx, desc, &fresh)) return e;
        const cn_camera camera = scene->camera;
        CN_HIP(hipDeviceSynchronize());  // waves in flight read the old buffers
        *scene = std::move(fresh);
        if (!(dirty & CN_DIRTY_CAMERA)) scene->camera = camera;
    } else if (dirty & (CN_DIRTY_MATERIALS | CN_DIRTY_LIGHTS | CN_DIRTY_ENVIRONMENT)) {
        // Everything fallible runs before the first change to the scene.
        if (!same_geometry(scene, desc)) {
            return fail(CN_E_SCENE, "the description's meshes or placements differ from the resident scene's "
                                    "(update with CN_DIRTY_GEOMETRY)");
        }
        if ((dirty & CN_DIRTY_ENVIRONMENT) && check_environment(desc->environment)) return CN_E_SCENE;
        if ((dirty & CN_DIRTY_LIGHTS) && check_delta_lights(desc)) return CN_E_SCENE;
        std::vector<InstanceRecord> records;
        std::vector<cn_material> materials;
        std::vector<uint32_t> tri_flags;
        std::vector<EmissiveTriangle> emitters;
        if (dirty & CN_DIRTY_MATERIALS) {
            MeshPool pool;
            if (int e = pool_meshes(desc, &pool)) return e;
            if (int e = prep_instances(desc, pool, &records, &materials, &tri_flags, &emitters, nullptr)) return e;
        }
        CN_HIP(hipDeviceSynchronize());  // waves in flight read the old buffers
        if (dirty & CN_DIRTY_MATERIALS) {
            std::string error;
            if (!patch_triangle_flags(scene->tris.ptr, uint32_t(scene->tris.count), tri_flags.data(),
                                      uint32_t(tri_flags.size()), &error))
                return fail(CN_E_DEVICE_FAULT, "%s", error.c_str());
            CN_HIP(scene->instances.upload(records.data(), records.size()));
            CN_HIP(scene->materials.upload(materials.data(), materials.size()));
            scene->host_records = std::move(records);
        } else {
            emitters = scene->host_emitters;
        }
        const bool lights = dirty & (CN_DIRTY_MATERIALS | CN_DIRTY_LIGHTS);
        if (lights) {
            std::vector<cn_delta_light> deltas = scene->host_deltas;
            if (dirty & CN_DIRTY_LIGHTS) deltas.assign(desc->delta_lights, desc->delta_lights + desc->delta_light_count);
            if (int e = install_lights(scene, std::move(emitters), deltas.data(), uint32_t(deltas.size()))) return e;
        }
        if (dirty & CN_DIRTY_ENVIRONMENT) {
            if (int e = install_environment(scene, desc->environment)) return e;
        }
        refresh_view(scene);
    }
    if (dirty & CN_DIRTY_CAMERA) scene->camera = desc->camera;
    // Captured launch graphs hold the old view: a new serial retires them.
    if (dirty & ~uint32_t(CN_DIRTY_CAMERA)) scene->serial = next_serial();
    scene->info.prep_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
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

int cn_scene_bvh8(cn_scene* scene, void* nodes, uint32_t node_capacity, uint32_t* node_count) {
    if (!scene) return fail(CN_E_INVALID, "cn_scene_bvh8: null scene");
    if (node_count) *node_count = uint32_t(scene->host_nodes8.size());
    if (nodes) {
        if (node_capacity < scene->host_nodes8.size()) return fail(CN_E_INVALID, "node buffer too small");
        std::memcpy(nodes, scene->host_nodes8.data(), scene->host_nodes8.size() * sizeof(BvhNode8Q));
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
