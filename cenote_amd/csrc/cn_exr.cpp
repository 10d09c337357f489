// cn_exr.cpp — OpenEXR output and input for the film (output.rs:37-193).
//
// Writes single-part scanline files, ZIP-compressed in 16-line blocks, with
// the ACEScg chromaticities attribute: `cn_write_exr` stores RGBA as f32;
// `cn_write_aov_exr` stores the beauty as f16 R/G/B/A, depth as f32 Z and
// the denoiser guides as f16 albedo.R/G/B and normal.X/Y/Z (the channel
// set of output.rs:62-136). Reads single-part scanline or single-level
// tiled files with NONE, RLE, ZIPS or ZIP compression and HALF, FLOAT or
// UINT channels, returning the first layer's R, G, B, A as f32 (a missing
// alpha reads as 1, output.rs:160-193).
//
// The file layout follows the OpenEXR file-format document (magic and
// version, attribute header, chunk offset table, chunks) and its ZIP/RLE
// pre-filters (byte interleave, then a byte-delta predictor).

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "cn_handles.h"

using namespace cn;

namespace {

// ACEScg primaries and white point (ACES S-2014-004; output.rs:17-23).
constexpr float ACESCG_CHROMATICITIES[8] = {0.713f, 0.293f, 0.165f, 0.830f, 0.128f, 0.044f, 0.32168f, 0.33767f};

enum PixelType : int32_t { PIX_UINT = 0, PIX_HALF = 1, PIX_FLOAT = 2 };
enum Compression : uint8_t { C_NONE = 0, C_RLE = 1, C_ZIPS = 2, C_ZIP = 3 };

// IEEE binary32 -> binary16, round to nearest even (overflow to inf, NaN kept).
uint16_t to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xff) return uint16_t(sign | 0x7c00u | (man ? (0x200u | (man >> 13)) : 0u));
    const int e = int(exp) - 127 + 15;
    if (e >= 31) return uint16_t(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return uint16_t(sign);
        man |= 0x800000u;
        const int shift = 14 - e;
        uint32_t half = man >> shift;
        const uint32_t rem = man & ((1u << shift) - 1u);
        const uint32_t mid = 1u << (shift - 1);
        if (rem > mid || (rem == mid && (half & 1u))) ++half;
        return uint16_t(sign | half);
    }
    uint32_t half = (uint32_t(e) << 10) | (man >> 13);
    const uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (half & 1u))) ++half;  // may carry into the exponent: correct
    return uint16_t(sign | half);
}

float from_half(uint16_t h) {
    const uint32_t sign = uint32_t(h & 0x8000u) << 16;
    const uint32_t exp = (h >> 10) & 0x1fu;
    uint32_t man = h & 0x3ffu;
    uint32_t x;
    if (exp == 0) {
        if (man == 0) {
            x = sign;
        } else {
            int e = -1;
            do {
                ++e;
                man <<= 1;
            } while (!(man & 0x400u));
            x = sign | (uint32_t(127 - 15 - e) << 23) | ((man & 0x3ffu) << 13);
        }
    } else if (exp == 31) {
        x = sign | 0x7f800000u | (man << 13);
    } else {
        x = sign | ((exp + 127 - 15) << 23) | (man << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

struct Out {
    std::vector<uint8_t> b;
    void u8(uint8_t v) { b.push_back(v); }
    void i32(int32_t v) { raw(&v, 4); }
    void u32(uint32_t v) { raw(&v, 4); }
    void u64(uint64_t v) { raw(&v, 8); }
    void f32(float v) { raw(&v, 4); }
    void str(const char* s) { raw(s, std::strlen(s) + 1); }
    void raw(const void* p, size_t n) {
        const uint8_t* c = static_cast<const uint8_t*>(p);
        b.insert(b.end(), c, c + n);
    }
    void attr(const char* name, const char* type, const std::vector<uint8_t>& value) {
        str(name);
        str(type);
        i32(int32_t(value.size()));
        raw(value.data(), value.size());
    }
};

// The ZIP/RLE pre-filter: bytes interleaved (even offsets first), then each
// byte replaced by its difference to the previous one, offset by 128.
void prefilter(const uint8_t* in, size_t n, std::vector<uint8_t>& out) {
    out.resize(n);
    size_t a = 0, b = (n + 1) / 2;
    for (size_t i = 0; i < n; ++i) out[(i & 1) ? b++ : a++] = in[i];
    int p = n ? out[0] : 0;
    for (size_t i = 1; i < n; ++i) {
        const int d = int(out[i]) - p + (128 + 256);
        p = out[i];
        out[i] = uint8_t(d);
    }
}

void unfilter(std::vector<uint8_t>& buf, std::vector<uint8_t>& out) {
    const size_t n = buf.size();
    for (size_t i = 1; i < n; ++i) buf[i] = uint8_t(int(buf[i - 1]) + int(buf[i]) - 128);
    out.resize(n);
    size_t a = 0, b = (n + 1) / 2;
    for (size_t i = 0; i < n; ++i) out[i] = (i & 1) ? buf[b++] : buf[a++];
}

struct ChannelOut {
    std::string name;
    int32_t type;
    const float* src;  // pixel p's value at src[p * stride]
    uint32_t stride;
};

int write_scanline_zip(const char* path, uint32_t width, uint32_t height, std::vector<ChannelOut> channels) {
    if (!path) return fail(CN_E_INVALID, "EXR: null path");
    if (width == 0 || height == 0 || width > (1u << 30) || height > (1u << 30))
        return fail(CN_E_INVALID, "EXR: bad size %ux%u", width, height);
    std::sort(channels.begin(), channels.end(),
              [](const ChannelOut& a, const ChannelOut& b) { return a.name < b.name; });  // the spec's order
    Out h;
    h.u32(20000630u);  // magic
    h.u32(2u);         // version 2, single-part scanline
    {
        Out v;
        for (const ChannelOut& c : channels) {
            v.str(c.name.c_str());
            v.i32(c.type);
            v.u8(0);  // pLinear
            v.u8(0);
            v.u8(0);
            v.u8(0);
            v.i32(1);
            v.i32(1);
        }
        v.u8(0);
        h.attr("channels", "chlist", v.b);
    }
    {
        Out v;
        for (float f : ACESCG_CHROMATICITIES) v.f32(f);
        h.attr("chromaticities", "chromaticities", v.b);
    }
    h.attr("compression", "compression", {C_ZIP});
    Out box;
    box.i32(0);
    box.i32(0);
    box.i32(int32_t(width) - 1);
    box.i32(int32_t(height) - 1);
    h.attr("dataWindow", "box2i", box.b);
    h.attr("displayWindow", "box2i", box.b);
    h.attr("lineOrder", "lineOrder", {0});  // increasing y
    {
        Out v;
        v.f32(1.0f);
        h.attr("pixelAspectRatio", "float", v.b);
    }
    {
        Out v;
        v.f32(0.0f);
        v.f32(0.0f);
        h.attr("screenWindowCenter", "v2f", v.b);
    }
    {
        Out v;
        v.f32(1.0f);
        h.attr("screenWindowWidth", "float", v.b);
    }
    h.u8(0);  // end of header

    constexpr uint32_t LINES = 16;  // ZIP's block height
    const uint32_t blocks = (height + LINES - 1) / LINES;
    const size_t table_at = h.b.size();
    for (uint32_t k = 0; k < blocks; ++k) h.u64(0);
    std::vector<uint8_t> raw, filtered, packed;
    for (uint32_t k = 0; k < blocks; ++k) {
        const uint32_t y0 = k * LINES, y1 = std::min(height, y0 + LINES);
        raw.clear();
        for (uint32_t y = y0; y < y1; ++y) {
            for (const ChannelOut& c : channels) {
                for (uint32_t x = 0; x < width; ++x) {
                    const float v = c.src[(size_t(y) * width + x) * c.stride];
                    if (c.type == PIX_HALF) {
                        const uint16_t hv = to_half(v);
                        raw.push_back(uint8_t(hv));
                        raw.push_back(uint8_t(hv >> 8));
                    } else {
                        uint8_t bytes[4];
                        std::memcpy(bytes, &v, 4);
                        raw.insert(raw.end(), bytes, bytes + 4);
                    }
                }
            }
        }
        prefilter(raw.data(), raw.size(), filtered);
        uLongf bound = compressBound(uLong(filtered.size()));
        packed.resize(bound);
        if (compress2(packed.data(), &bound, filtered.data(), uLong(filtered.size()), Z_DEFAULT_COMPRESSION) != Z_OK)
            return fail(CN_E_INVALID, "EXR: zlib failed");
        packed.resize(bound);
        const uint64_t at = h.b.size();
        std::memcpy(&h.b[table_at + 8 * k], &at, 8);
        h.i32(int32_t(y0));
        // A block that does not shrink is stored raw (its size then equals
        // the uncompressed size, which is how readers tell).
        const std::vector<uint8_t>& body = packed.size() < raw.size() ? packed : raw;
        h.i32(int32_t(body.size()));
        h.raw(body.data(), body.size());
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(CN_E_INVALID, "EXR: cannot open %s for writing", path);
    const size_t wrote = std::fwrite(h.b.data(), 1, h.b.size(), f);
    const int closed = std::fclose(f);
    if (wrote != h.b.size() || closed != 0) return fail(CN_E_INVALID, "EXR: short write to %s", path);
    return CN_OK;
}

// --- reading ---------------------------------------------------------------------

struct In {
    const uint8_t* p;
    size_t n, at = 0;
    bool ok = true;
    bool need(size_t k) {
        if (at + k > n) ok = false;
        return ok;
    }
    template <typename T>
    T get() {
        T v{};
        if (need(sizeof(T))) {
            std::memcpy(&v, p + at, sizeof(T));
            at += sizeof(T);
        }
        return v;
    }
    std::string str() {
        std::string s;
        while (need(1) && p[at] != 0) s.push_back(char(p[at++]));
        if (need(1)) ++at;
        return s;
    }
};

struct ChannelIn {
    std::string name;
    int32_t type;
};

size_t type_bytes(int32_t t) { return t == PIX_HALF ? 2 : 4; }

bool rle_decode(const uint8_t* in, size_t n, std::vector<uint8_t>& out, size_t expect) {
    out.clear();
    size_t i = 0;
    while (i < n) {
        const int8_t c = int8_t(in[i++]);
        if (c < 0) {
            const size_t count = size_t(-int(c));
            if (i + count > n) return false;
            out.insert(out.end(), in + i, in + i + count);
            i += count;
        } else {
            if (i >= n) return false;
            out.insert(out.end(), size_t(c) + 1, in[i++]);
        }
        if (out.size() > expect) return false;
    }
    return out.size() == expect;
}

// Decompress one chunk's data into `raw` (expect bytes).
int decode_chunk(uint8_t comp, const uint8_t* data, size_t size, size_t expect, std::vector<uint8_t>& raw) {
    if (size == expect || comp == C_NONE) {
        if (size != expect) return fail(CN_E_INVALID, "EXR: uncompressed chunk of %zu bytes, expected %zu", size, expect);
        raw.assign(data, data + size);
        return CN_OK;
    }
    std::vector<uint8_t> filtered;
    if (comp == C_RLE) {
        if (!rle_decode(data, size, filtered, expect)) return fail(CN_E_INVALID, "EXR: bad RLE chunk");
    } else if (comp == C_ZIP || comp == C_ZIPS) {
        filtered.resize(expect);
        uLongf len = uLongf(expect);
        if (uncompress(filtered.data(), &len, data, uLong(size)) != Z_OK || len != expect)
            return fail(CN_E_INVALID, "EXR: bad ZIP chunk");
    } else {
        return fail(CN_E_INVALID, "EXR: compression %u is not supported", comp);
    }
    unfilter(filtered, raw);
    return CN_OK;
}

float sample_at(const uint8_t* p, int32_t type) {
    if (type == PIX_HALF) {
        uint16_t h;
        std::memcpy(&h, p, 2);
        return from_half(h);
    }
    if (type == PIX_FLOAT) {
        float f;
        std::memcpy(&f, p, 4);
        return f;
    }
    uint32_t u;
    std::memcpy(&u, p, 4);
    return float(u);
}

int read_rgba(const std::vector<uint8_t>& file, uint32_t* width_out, uint32_t* height_out, float* rgba,
              size_t capacity) {
    In in{file.data(), file.size()};
    if (in.get<uint32_t>() != 20000630u) return fail(CN_E_INVALID, "EXR: bad magic number");
    const uint32_t version = in.get<uint32_t>();
    if ((version & 0xff) != 2) return fail(CN_E_INVALID, "EXR: version %u is not supported", version & 0xff);
    if (version & 0x1800u) return fail(CN_E_INVALID, "EXR: deep and multi-part files are not supported");
    const bool tiled = version & 0x200u;
    std::vector<ChannelIn> channels;
    uint8_t comp = 0xff;
    int32_t box[4] = {0, 0, -1, -1};
    uint32_t tile_w = 0, tile_h = 0, tile_mode = 0;
    while (in.ok) {
        const std::string name = in.str();
        if (name.empty()) break;
        const std::string type = in.str();
        const int32_t size = in.get<int32_t>();
        if (size < 0 || !in.need(size_t(size))) break;
        const size_t end = in.at + size_t(size);
        if (name == "channels") {
            while (in.at < end) {
                const std::string cname = in.str();
                if (cname.empty()) break;
                ChannelIn c{cname, in.get<int32_t>()};
                in.get<uint32_t>();  // pLinear + reserved
                const int32_t xs = in.get<int32_t>(), ys = in.get<int32_t>();
                if (xs != 1 || ys != 1) return fail(CN_E_INVALID, "EXR: subsampled channels are not supported");
                if (c.type < 0 || c.type > 2) return fail(CN_E_INVALID, "EXR: bad pixel type");
                channels.push_back(c);
            }
        } else if (name == "compression") {
            comp = in.get<uint8_t>();
        } else if (name == "dataWindow") {
            for (int k = 0; k < 4; ++k) box[k] = in.get<int32_t>();
        } else if (name == "tiles") {
            tile_w = in.get<uint32_t>();
            tile_h = in.get<uint32_t>();
            tile_mode = in.get<uint8_t>();
        }
        in.at = end;
    }
    if (!in.ok) return fail(CN_E_INVALID, "EXR: truncated header");
    if (channels.empty() || comp == 0xff) return fail(CN_E_INVALID, "EXR: header lacks channels or compression");
    const int64_t w64 = int64_t(box[2]) - box[0] + 1, h64 = int64_t(box[3]) - box[1] + 1;
    if (w64 <= 0 || h64 <= 0 || w64 > (1 << 30) || h64 > (1 << 30)) return fail(CN_E_INVALID, "EXR: bad data window");
    const uint32_t width = uint32_t(w64), height = uint32_t(h64);
    *width_out = width;
    *height_out = height;
    if (!rgba) return CN_OK;
    if (capacity < size_t(width) * height * 4) return fail(CN_E_INVALID, "EXR: output buffer too small");
    int slot[4] = {-1, -1, -1, -1};  // R, G, B, A channel indices
    for (size_t k = 0; k < channels.size(); ++k) {
        const std::string& n = channels[k].name;
        if (n == "R") slot[0] = int(k);
        if (n == "G") slot[1] = int(k);
        if (n == "B") slot[2] = int(k);
        if (n == "A") slot[3] = int(k);
    }
    for (size_t i = 0; i < size_t(width) * height; ++i) {
        rgba[4 * i] = rgba[4 * i + 1] = rgba[4 * i + 2] = 0.0f;
        rgba[4 * i + 3] = 1.0f;
    }
    size_t pixel_bytes = 0;
    for (const ChannelIn& c : channels) pixel_bytes += type_bytes(c.type);
    std::vector<uint8_t> raw;
    // Store the lines [y0, y0 + lines) x [x0, x0 + span) decoded from `raw`.
    auto store = [&](int32_t x0, int32_t y0, uint32_t span, uint32_t lines) {
        size_t off = 0;
        for (uint32_t ly = 0; ly < lines; ++ly) {
            for (size_t c = 0; c < channels.size(); ++c) {
                const size_t tb = type_bytes(channels[c].type);
                int dest = -1;
                for (int k = 0; k < 4; ++k)
                    if (slot[k] == int(c)) dest = k;
                if (dest >= 0) {
                    for (uint32_t x = 0; x < span; ++x) {
                        const size_t px = size_t(y0 - box[1] + int32_t(ly)) * width + size_t(x0 - box[0]) + x;
                        rgba[4 * px + dest] = sample_at(raw.data() + off + x * tb, channels[c].type);
                    }
                }
                off += span * tb;
            }
        }
    };
    if (!tiled) {
        const uint32_t lines = (comp == C_ZIP) ? 16u : 1u;
        const uint32_t blocks = (height + lines - 1) / lines;
        std::vector<uint64_t> table(blocks);
        for (auto& t : table) t = in.get<uint64_t>();
        for (uint32_t k = 0; k < blocks && in.ok; ++k) {
            In c{file.data(), file.size(), size_t(table[k])};
            const int32_t y = c.get<int32_t>();
            const int32_t size = c.get<int32_t>();
            if (!c.ok || size < 0 || !c.need(size_t(size))) return fail(CN_E_INVALID, "EXR: bad chunk");
            if (y < box[1] || y > box[3]) return fail(CN_E_INVALID, "EXR: chunk outside the data window");
            const uint32_t n = std::min<uint32_t>(lines, uint32_t(box[3] - y + 1));
            if (int e = decode_chunk(comp, file.data() + c.at, size_t(size), size_t(n) * width * pixel_bytes, raw))
                return e;
            store(box[0], y, width, n);
        }
    } else {
        if ((tile_mode & 0xf) != 0) return fail(CN_E_INVALID, "EXR: only single-level tiled files are supported");
        if (tile_w == 0 || tile_h == 0) return fail(CN_E_INVALID, "EXR: bad tile size");
        const uint32_t tx = (width + tile_w - 1) / tile_w, ty = (height + tile_h - 1) / tile_h;
        std::vector<uint64_t> table(size_t(tx) * ty);
        for (auto& t : table) t = in.get<uint64_t>();
        for (size_t k = 0; k < table.size() && in.ok; ++k) {
            In c{file.data(), file.size(), size_t(table[k])};
            const int32_t dx = c.get<int32_t>(), dy = c.get<int32_t>();
            c.get<int32_t>();  // level x
            c.get<int32_t>();  // level y
            const int32_t size = c.get<int32_t>();
            if (!c.ok || size < 0 || !c.need(size_t(size))) return fail(CN_E_INVALID, "EXR: bad tile");
            if (dx < 0 || dy < 0 || uint32_t(dx) >= tx || uint32_t(dy) >= ty)
                return fail(CN_E_INVALID, "EXR: tile outside the data window");
            const uint32_t x0 = uint32_t(dx) * tile_w, y0 = uint32_t(dy) * tile_h;
            const uint32_t span = std::min(tile_w, width - x0), lines = std::min(tile_h, height - y0);
            if (int e = decode_chunk(comp, file.data() + c.at, size_t(size), size_t(lines) * span * pixel_bytes, raw))
                return e;
            store(box[0] + int32_t(x0), box[1] + int32_t(y0), span, lines);
        }
    }
    if (!in.ok) return fail(CN_E_INVALID, "EXR: truncated offset table");
    return CN_OK;
}

}  // namespace

extern "C" {

int cn_write_exr(const char* path, uint32_t width, uint32_t height, const float* rgba) {
    if (!rgba) return fail(CN_E_INVALID, "cn_write_exr: null pixels");
    return write_scanline_zip(path, width, height,
                              {{"R", PIX_FLOAT, rgba, 4}, {"G", PIX_FLOAT, rgba + 1, 4}, {"B", PIX_FLOAT, rgba + 2, 4},
                               {"A", PIX_FLOAT, rgba + 3, 4}});
}

int cn_write_aov_exr(const char* path, uint32_t width, uint32_t height, const float* beauty, const float* albedo,
                     const float* normal, const float* depth) {
    if (!beauty || !albedo || !normal || !depth) return fail(CN_E_INVALID, "cn_write_aov_exr: null layer");
    return write_scanline_zip(path, width, height,
                              {{"R", PIX_HALF, beauty, 4},
                               {"G", PIX_HALF, beauty + 1, 4},
                               {"B", PIX_HALF, beauty + 2, 4},
                               {"A", PIX_HALF, beauty + 3, 4},
                               {"Z", PIX_FLOAT, depth, 1},
                               {"albedo.R", PIX_HALF, albedo, 4},
                               {"albedo.G", PIX_HALF, albedo + 1, 4},
                               {"albedo.B", PIX_HALF, albedo + 2, 4},
                               {"normal.X", PIX_HALF, normal, 4},
                               {"normal.Y", PIX_HALF, normal + 1, 4},
                               {"normal.Z", PIX_HALF, normal + 2, 4}});
}

int cn_read_exr(const char* path, uint32_t* width, uint32_t* height, float* rgba, size_t capacity) {
    if (!path || !width || !height) return fail(CN_E_INVALID, "cn_read_exr: null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CN_E_INVALID, "EXR: cannot open %s", path);
    std::vector<uint8_t> file;
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) file.insert(file.end(), buf, buf + got);
    std::fclose(f);
    return read_rgba(file, width, height, rgba, capacity);
}

}  // extern "C"
