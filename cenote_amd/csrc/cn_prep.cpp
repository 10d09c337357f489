// cn_prep.cpp — host-side scene preparation (see cn_prep.h).

#include "cn_prep.h"

#include <algorithm>
#include <cmath>
#include <thread>

namespace cn {

namespace {

constexpr uint32_t KIND_TRIANGLE = LIGHT_KIND_TRIANGLE;
constexpr uint32_t KIND_DISTANT = LIGHT_KIND_DISTANT;
constexpr uint32_t KIND_POINT = LIGHT_KIND_POINT;

v3 rec_anchor(const cn_light_record& r) { return V3(r.anchor[0], r.anchor[1], r.anchor[2]); }
v3 rec_edge1(const cn_light_record& r) { return V3(r.edge1[0], r.edge1[1], r.edge1[2]); }
v3 rec_edge2(const cn_light_record& r) { return V3(r.edge2[0], r.edge2[1], r.edge2[2]); }
v3 rec_emission(const cn_light_record& r) { return V3(r.emission[0], r.emission[1], r.emission[2]); }

void put3(float* dst, v3 v) {
    dst[0] = v.x;
    dst[1] = v.y;
    dst[2] = v.z;
}

// lights.rs:135-137: half the length of the edges' cross product, the
// length taken in f32 (glam) and the halving in f64.
double record_area(const cn_light_record& r) {
    return double(length(cross(rec_edge1(r), rec_edge2(r)))) / 2.0;
}

// lights.rs:123-130
double record_power(const cn_light_record& r) {
    const double luma = double(luminance(rec_emission(r)));
    if (r.kind == KIND_TRIANGLE) return M_PI * luma * record_area(r);
    if (r.kind == KIND_DISTANT) return luma;
    return 4.0 * M_PI * luma;
}

cn_light_record fresh_record(uint32_t kind, v3 anchor, v3 e1, v3 e2, v3 emission, uint32_t instance,
                             uint32_t primitive) {
    cn_light_record r{};
    put3(r.anchor, anchor);
    r.pdf = 0.0f;
    put3(r.edge1, e1);
    r.alias_threshold = 1.0f;
    put3(r.edge2, e2);
    r.alias_index = 0;
    put3(r.emission, emission);
    r.instance = instance;
    r.kind = kind;
    r.primitive = primitive;
    return r;
}

}  // namespace

float luminance(v3 c) { return c.x * 0.2722287f + c.y * 0.6740818f + c.z * 0.0536895f; }

uint32_t material_class(const cn_material& m) {
    return (m.metalness > 0.0f ? 1u : 0u) | (m.metalness < 1.0f ? 2u : 0u) | (m.transmission_weight > 0.0f ? 4u : 0u) |
           (m.coat_weight > 0.0f ? 8u : 0u) | (m.fuzz_weight > 0.0f ? 16u : 0u);
}

int32_t log2_transmittance_fixed(float opacity) {
    const double transmittance = 1.0 - double(opacity);
    if (!(transmittance > 0.0)) return LOG2_FIXED_FLOOR;
    if (transmittance >= 1.0) return 0;
    const double units = std::nearbyint(std::log2(transmittance) * double(1 << LOG2_FIXED_SHIFT));
    return units < double(LOG2_FIXED_FLOOR) ? LOG2_FIXED_FLOOR : int32_t(units);
}

std::vector<cn_light_record> build_light_table(const std::vector<EmissiveTriangle>& tris,
                                               const cn_delta_light* deltas, uint32_t delta_count,
                                               double* total_power) {
    // lights.rs:236-275 — triangles first in input order, then deltas.
    std::vector<cn_light_record> records;
    records.reserve(tris.size() + delta_count);
    for (const EmissiveTriangle& t : tris) {
        records.push_back(fresh_record(KIND_TRIANGLE, t.corners[0], t.corners[1] - t.corners[0],
                                       t.corners[2] - t.corners[0], t.emission, t.instance, t.primitive));
    }
    for (uint32_t i = 0; i < delta_count; ++i) {
        const cn_delta_light& d = deltas[i];
        records.push_back(fresh_record(d.kind == CN_LIGHT_DISTANT ? KIND_DISTANT : KIND_POINT,
                                       V3(d.vector[0], d.vector[1], d.vector[2]), S3(0.0f), S3(0.0f),
                                       V3(d.color[0], d.color[1], d.color[2]), LIGHT_NONE, 0));
    }

    const size_t n = records.size();
    std::vector<double> power(n);
    for (size_t i = 0; i < n; ++i) power[i] = record_power(records[i]);
    double total = 0.0;
    for (size_t i = 0; i < n; ++i) total += power[i];  // a left fold, like Iterator::sum
    if (total_power) *total_power = total;
    if (total <= 0.0) return records;

    for (size_t i = 0; i < n; ++i) {
        const double selection = power[i] / total;
        if (records[i].kind == KIND_TRIANGLE) {
            const double area = record_area(records[i]);
            records[i].pdf = area > 0.0 ? float(selection / area) : 0.0f;
        } else {
            records[i].pdf = float(selection);
        }
    }

    // Vose's alias method with the reference's worklist discipline: both
    // lists keep index order, and each round takes the last of each.
    std::vector<double> share(n);
    const double count = double(n);
    for (size_t i = 0; i < n; ++i) share[i] = power[i] * count / total;
    std::vector<uint32_t> poor_list, rich_list;
    poor_list.reserve(n);
    rich_list.reserve(n);
    for (uint32_t i = 0; i < n; ++i) (share[i] < 1.0 ? poor_list : rich_list).push_back(i);
    while (!poor_list.empty() && !rich_list.empty()) {
        const uint32_t poor = poor_list.back();
        poor_list.pop_back();
        const uint32_t rich = rich_list.back();
        rich_list.pop_back();
        records[poor].alias_threshold = float(share[poor]);
        records[poor].alias_index = rich;
        share[rich] -= 1.0 - share[poor];
        (share[rich] < 1.0 ? poor_list : rich_list).push_back(rich);
    }
    return records;
}

EnvironmentTables build_environment_tables(uint32_t width, uint32_t height, const float* texels) {
    const size_t w = width, h = height;
    EnvironmentTables t;
    std::vector<double> lum(w * h);
    for (size_t i = 0; i < w * h; ++i) {
        lum[i] = double(luminance(V3(texels[4 * i], texels[4 * i + 1], texels[4 * i + 2])));
    }

    // Dilation: the 3x3 maximum (wrapping u, clamping v) is exact in any
    // order, so it runs as a horizontal then a vertical pass, rows in
    // parallel; fmax ignores a NaN the way Rust's f64::max does. The sums
    // below keep the reference's sequential order.
    std::vector<double> hmax(w * h), weight(w * h);
    auto rows_parallel = [&](auto&& body) {
        const unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (unsigned id = 0; id < threads; ++id) {
            pool.emplace_back([&, id]() {
                for (size_t row = id; row < h; row += threads) body(row);
            });
        }
        for (auto& th : pool) th.join();
    };
    rows_parallel([&](size_t row) {
        const double* l = &lum[row * w];
        for (size_t col = 0; col < w; ++col) {
            const size_t left = col == 0 ? w - 1 : col - 1;
            const size_t right = col + 1 == w ? 0 : col + 1;
            hmax[row * w + col] = std::fmax(std::fmax(l[left], l[col]), l[right]);
        }
    });
    rows_parallel([&](size_t row) {
        const size_t up = row == 0 ? 0 : row - 1;
        const size_t down = std::min(row + 1, h - 1);
        const double sin_theta = std::sin((double(row) + 0.5) / double(h) * M_PI);
        for (size_t col = 0; col < w; ++col) {
            const double peak =
                std::fmax(std::fmax(std::fmax(0.0, hmax[up * w + col]), hmax[row * w + col]), hmax[down * w + col]);
            weight[row * w + col] = peak * sin_theta;
        }
    });

    double total = 0.0;
    for (size_t i = 0; i < w * h; ++i) total += weight[i];  // environment.rs:157, one flat fold

    std::vector<double> row_sum(h);
    rows_parallel([&](size_t row) {
        double acc = 0.0;
        for (size_t col = 0; col < w; ++col) acc += weight[row * w + col];
        row_sum[row] = acc;
    });

    t.marginal.assign(h + 1, 0.0f);
    double running = 0.0;
    for (size_t row = 0; row < h; ++row) {
        running += row_sum[row];
        t.marginal[row + 1] = total > 0.0 ? float(running / total) : 0.0f;
    }
    t.marginal[h] = 1.0f;

    t.conditional.assign(h * (w + 1), 0.0f);
    rows_parallel([&](size_t row) {
        float* cdf = &t.conditional[row * (w + 1)];
        double acc = 0.0;
        for (size_t col = 0; col < w; ++col) {
            acc += weight[row * w + col];
            cdf[col + 1] = row_sum[row] > 0.0 ? float(acc / row_sum[row]) : 0.0f;
        }
        cdf[w] = 1.0f;
    });

    const double scale = total > 0.0 ? double(w * h) / total : 0.0;
    t.pdfs.resize(w * h);
    for (size_t i = 0; i < w * h; ++i) t.pdfs[i] = float(weight[i] * scale);

    double power = 0.0;
    for (size_t row = 0; row < h; ++row) {
        const double theta0 = double(row) / double(h) * M_PI;
        const double theta1 = double(row + 1) / double(h) * M_PI;
        const double solid = 2.0 * M_PI / double(w) * (std::cos(theta0) - std::cos(theta1));
        double acc = 0.0;
        for (size_t col = 0; col < w; ++col) acc += lum[row * w + col];
        power += solid * acc;
    }
    t.power = power;

    // A table that does not rise monotonically (non-finite texels) gets one
    // bucket spanning every index, which is the plain search.
    auto guides = [&](const float* cdfs, size_t stride, uint32_t count, size_t tables, uint32_t* buckets,
                      std::vector<uint32_t>* out) {
        *buckets = guide_buckets(count);
        out->assign(tables * (*buckets + 1), 0u);
        std::vector<char> ok(tables, 1);
        auto one = [&](size_t k) {
            ok[k] = build_cdf_guide(cdfs + k * stride, count, *buckets, out->data() + k * (*buckets + 1));
        };
        if (tables == 1) one(0);
        else rows_parallel(one);
        if (std::find(ok.begin(), ok.end(), 0) != ok.end()) {
            *buckets = 1;
            out->assign(tables * 2, 0u);
            for (size_t k = 0; k < tables; ++k) (*out)[2 * k + 1] = count - 1;
        }
    };
    guides(t.marginal.data(), 0, uint32_t(h), 1, &t.marginal_buckets, &t.marginal_guide);
    guides(t.conditional.data(), w + 1, uint32_t(w), h, &t.conditional_buckets, &t.conditional_guide);
    return t;
}

uint32_t guide_buckets(uint32_t count) {
    uint32_t b = 1;
    while (b < count && b < (1u << 20)) b <<= 1;
    return b;
}

static uint32_t cdf_index(const float* cdf, uint32_t count, float u) {
    uint32_t lo = 0, hi = count;
    while (lo + 1 < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid;
    }
    return lo;
}

bool build_cdf_guide(const float* cdf, uint32_t count, uint32_t buckets, uint32_t* guide) {
    bool monotone = cdf[0] <= 0.0f;
    for (uint32_t i = 1; i < count && monotone; ++i) monotone = cdf[i - 1] <= cdf[i];
    if (!monotone) return false;
    for (uint32_t g = 0; g <= buckets; ++g) guide[g] = cdf_index(cdf, count, float(g) / float(buckets));
    return true;
}

float select_probability(double env_power, double light_power) {
    if (light_power == 0.0) return env_power > 0.0 ? 1.0f : 0.0f;
    return float(env_power / (env_power + light_power));
}

v3 transform_point(const float r[12], v3 p) {
    return V3(((r[0] * p.x + r[1] * p.y) + r[2] * p.z) + r[3], ((r[4] * p.x + r[5] * p.y) + r[6] * p.z) + r[7],
              ((r[8] * p.x + r[9] * p.y) + r[10] * p.z) + r[11]);
}

bool invert_affine(const float m[12], float out[12]) {
    for (int i = 0; i < 12; ++i) {
        if (!std::isfinite(m[i])) return false;
    }
    const double a00 = m[0], a01 = m[1], a02 = m[2], a10 = m[4], a11 = m[5], a12 = m[6], a20 = m[8], a21 = m[9],
                 a22 = m[10];
    const double c00 = a11 * a22 - a12 * a21;
    const double c01 = a12 * a20 - a10 * a22;
    const double c02 = a10 * a21 - a11 * a20;
    const double det = a00 * c00 + a01 * c01 + a02 * c02;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
    double inv[3][3];
    inv[0][0] = c00 / det;
    inv[1][0] = c01 / det;
    inv[2][0] = c02 / det;
    inv[0][1] = (a02 * a21 - a01 * a22) / det;
    inv[1][1] = (a00 * a22 - a02 * a20) / det;
    inv[2][1] = (a01 * a20 - a00 * a21) / det;
    inv[0][2] = (a01 * a12 - a02 * a11) / det;
    inv[1][2] = (a02 * a10 - a00 * a12) / det;
    inv[2][2] = (a00 * a11 - a01 * a10) / det;
    const double t[3] = {m[3], m[7], m[11]};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) out[4 * i + j] = float(inv[i][j]);
        out[4 * i + 3] = float(-(inv[i][0] * t[0] + inv[i][1] * t[1] + inv[i][2] * t[2]));
    }
    for (int i = 0; i < 12; ++i) {
        if (!std::isfinite(out[i])) return false;
    }
    return true;
}

bool camera_view(const cn_camera& c, float aspect, CameraView* out) {
    const v3 position = V3(c.position[0], c.position[1], c.position[2]);
    const v3 look_at = V3(c.look_at[0], c.look_at[1], c.look_at[2]);
    const v3 up_hint = V3(c.up[0], c.up[1], c.up[2]);
    const v3 forward = normalize(look_at - position);
    const v3 right = normalize(cross(forward, up_hint));
    if (!std::isfinite(forward.x + forward.y + forward.z) || !std::isfinite(right.x + right.y + right.z)) {
        return false;
    }
    const v3 up = cross(right, forward);
    const float rads = 3.14159265358979323846f / 180.0f;  // f32::to_radians's factor
    const float half_height = std::tan(c.vfov_degrees * rads / 2.0f);
    v3 r = (right * half_height) * aspect;
    v3 u = up * half_height;
    v3 f = forward;
    float aperture = 0.0f;
    if (c.aperture_radius > 0.0f) {
        const float focus = c.focus_distance > 0.0f ? c.focus_distance : length(position - look_at);
        r = r * focus;
        u = u * focus;
        f = f * focus;
        aperture = c.aperture_radius;
    }
    out->position = make_float3(position.x, position.y, position.z);
    out->right = make_float3(r.x, r.y, r.z);
    out->up = make_float3(u.x, u.y, u.z);
    out->forward = make_float3(f.x, f.y, f.z);
    out->aperture = aperture;
    return true;
}

}  // namespace cn
