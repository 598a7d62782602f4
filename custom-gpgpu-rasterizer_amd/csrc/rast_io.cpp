// rast_io.cpp -- OBJ meshes in, BMP framebuffers out (host only).
//
// OBJ: the parse of CCR/ObjParser.cpp (kdedeurw/Custom-GPGPU-Rasterizer),
// restated for a text buffer: the same line classification, field splitting,
// float/int conversion (strtof / strtol on the same characters as std::stof
// / std::stoi), vertex de-duplication and tangent accumulation, all in
// binary32 in the reference's operation order.  The de-duplication keeps the
// reference's semantics (with uvs and normals: the last earlier vertex with
// the same position index, the same uv index and a normal equal within
// 2 ulp; otherwise the first with the same indices) but finds candidates
// through a hash of the index pair instead of a scan of all earlier vertices.
#include "rast_io.h"

#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <unordered_map>
#include <vector>

#include <zlib.h>

namespace {

thread_local std::string g_err;

int fail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return RAST_E_INVALID_ARG;
}

struct V2 {
    float x, y;
};
struct V3 {
    float x, y, z;
};

// AreEqual(a, b, 2) (CCR/MathUtilities.h:30-38), binary32
bool are_equal(float a, float b)
{
    const float eps = std::numeric_limits<float>::epsilon();
    return std::fabs(a - b) <= eps * std::fabs(a + b) * 2.f || std::fabs(a - b) < std::numeric_limits<float>::min();
}
bool same_normal(const V3 &a, const V3 &b) { return are_equal(a.x, b.x) && are_equal(a.y, b.y) && are_equal(a.z, b.z); }

float dot3(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Normalize (CCR/Vector.h:41-56): m = sqrt(Dot(v, v)); zero stays zero; v *= 1 / m
V3 normalized(V3 v)
{
    const float m = std::sqrt(dot3(v, v));
    if (m == 0.f) return V3{ 0.f, 0.f, 0.f };
    const float inv = 1.f / m;
    return V3{ v.x * inv, v.y * inv, v.z * inv };
}

// Reject(v, t) = v - t * (Dot(v, t) / Dot(t, t)) (CCR/Vector.h:66-74)
V3 reject(const V3 &v, const V3 &t)
{
    const float s = dot3(v, t) / dot3(t, t);
    return V3{ v.x - t.x * s, v.y - t.y * s, v.z - t.z * s };
}

// std::getline(stream, token, ' ') three times (GetFirstSecondThird,
// CCR/ObjParser.cpp:180-192)
void first_three(const std::string &s, std::string out[3])
{
    size_t pos = 0;
    for (int i = 0; i < 3; ++i) {
        if (pos >= s.size()) {
            out[i].clear();
            continue;
        }
        const size_t sp = s.find(' ', pos);
        out[i] = s.substr(pos, sp == std::string::npos ? std::string::npos : sp - pos);
        pos = sp == std::string::npos ? s.size() : sp + 1;
    }
}

// std::stof: leading whitespace, then the longest number prefix; none -> error
bool to_float(const std::string &s, float &out)
{
    const char *b = s.c_str();
    char *e = nullptr;
    errno = 0;
    const float v = std::strtof(b, &e);
    if (e == b) return false;
    out = v;
    return true;
}

bool to_int(const std::string &s, long &out)
{
    const char *b = s.c_str();
    char *e = nullptr;
    errno = 0;
    const long v = std::strtol(b, &e, 10);
    if (e == b || errno) return false;
    out = v;
    return true;
}

struct Parsed {
    std::vector<V3> pos, nrm;
    std::vector<V2> uv;
    std::vector<long> pi, ti, ni;  // per face vertex: position, uv, normal index (0-based)
};

int parse(const char *text, size_t len, bool invert_y, Parsed &P)
{
    size_t at = 0, lineno = 0;
    std::string line, tok[3];
    while (at < len) {
        size_t nl = at;
        while (nl < len && text[nl] != '\n') ++nl;
        line.assign(text + at, nl - at);
        at = nl + 1;
        ++lineno;
        if (line.empty() || line[0] == '#') continue;
        const std::string prefix = line.substr(0, line.find(' '));
        if (prefix == "v") {  // "v" + two spaces (CCR/ObjParser.cpp:43-47)
            first_three(line.size() > 3 ? line.substr(3) : std::string(), tok);
            V3 p;
            if (!to_float(tok[0], p.x) || !to_float(tok[1], p.y) || !to_float(tok[2], p.z))
                return fail("line %zu: bad position", lineno);
            P.pos.push_back(p);
        } else if (prefix == "f") {  // StoreFace (CCR/ObjParser.cpp:118-171)
            first_three(line.size() > 2 ? line.substr(2) : std::string(), tok);
            std::vector<long> idx;
            for (int i = 0; i < 3; ++i) {
                if (tok[i].empty()) return fail("line %zu: face with fewer than three vertices", lineno);
                size_t from = 0;
                while (true) {
                    const size_t slash = tok[i].find('/', from);
                    const std::string field = tok[i].substr(from, slash == std::string::npos ? std::string::npos : slash - from);
                    long v;
                    if (!to_int(field, v)) return fail("line %zu: empty or bad index field", lineno);
                    if (v < 1) return fail("line %zu: index %ld (relative and zero indices are not supported)", lineno, v);
                    idx.push_back(v - 1);
                    if (slash == std::string::npos) break;
                    from = slash + 1;
                }
            }
            const size_t k = idx.size() / 3;
            if (k == 0 || idx.size() % 3 != 0) return fail("line %zu: vertices with different index counts", lineno);
            for (size_t q = 0; q < idx.size(); ++q) {
                const size_t c = q % k;
                if (c == 0) P.pi.push_back(idx[q]);
                else if (c == 1) P.ti.push_back(idx[q]);
                else if (c == 2) P.ni.push_back(idx[q]);
            }
        } else if (prefix == "vt") {  // StoreUV (CCR/ObjParser.cpp:173-179)
            first_three(line.size() > 3 ? line.substr(3) : std::string(), tok);
            V2 t;
            if (!to_float(tok[0], t.x) || !to_float(tok[1], t.y)) return fail("line %zu: bad uv", lineno);
            if (invert_y) t.y = 1 - t.y;
            P.uv.push_back(t);
        } else if (prefix == "vn") {  // StoreNormal (CCR/ObjParser.cpp:166-171)
            first_three(line.size() > 3 ? line.substr(3) : std::string(), tok);
            V3 n;
            if (!to_float(tok[0], n.x) || !to_float(tok[1], n.y) || !to_float(tok[2], n.z))
                return fail("line %zu: bad normal", lineno);
            P.nrm.push_back(normalized(n));
        }
        // "o", "g" and unknown prefixes are reported by the reference and skipped
    }
    return RAST_OK;
}

struct Out {
    std::vector<rast_ivertex> v;
    std::vector<uint32_t> idx;
    int type = 0;
};

rast_ivertex make_vertex(const V3 &p, const V2 &uv, const V3 &n)
{
    rast_ivertex v;
    v.p[0] = p.x; v.p[1] = p.y; v.p[2] = p.z;
    v.uv[0] = uv.x; v.uv[1] = uv.y;
    v.n[0] = n.x; v.n[1] = n.y; v.n[2] = n.z;
    v.tan[0] = v.tan[1] = v.tan[2] = 0.f;
    v.c[0] = v.c[1] = v.c[2] = 1.f;
    return v;
}

// AssignVertices, vector overload (CCR/ObjParser.cpp:381-480)
int assign(const Parsed &P, Out &O)
{
    const size_t n = P.pi.size();
    const bool has_uv = !P.uv.empty(), has_n = !P.nrm.empty();
    if (has_uv) O.type |= RIO_VERTEX_UV;
    if (has_n) O.type |= RIO_VERTEX_NORMAL;
    if (has_n && !has_uv) return fail("normals without uvs are not supported (the reference reads past its uv list)");
    if (has_uv && P.ti.size() != n) return fail("faces without uv indices in a file with uvs");
    if (has_n && P.ni.size() != n) return fail("faces without normal indices in a file with normals");
    for (size_t i = 0; i < n; ++i) {
        if ((size_t)P.pi[i] >= P.pos.size()) return fail("position index %ld out of range", P.pi[i] + 1);
        if (has_uv && (size_t)P.ti[i] >= P.uv.size()) return fail("uv index %ld out of range", P.ti[i] + 1);
        if (has_n && (size_t)P.ni[i] >= P.nrm.size()) return fail("normal index %ld out of range", P.ni[i] + 1);
    }
    // earlier unique vertices by (position index, uv index): (normal index, vertex)
    struct Seen {
        long ni;
        uint32_t vertex;
    };
    std::unordered_map<uint64_t, std::vector<Seen>> seen;
    O.v.reserve(n);
    O.idx.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        const long ti = has_uv ? P.ti[i] : 0;
        const uint64_t key = (uint64_t)P.pi[i] << 32 | (uint64_t)(uint32_t)ti;
        std::vector<Seen> &cand = seen[key];
        bool found = false;
        uint32_t vi = 0;
        if (has_uv && has_n) {
            for (const Seen &s : cand)  // no break: the last match wins
                if (same_normal(P.nrm[(size_t)s.ni], P.nrm[(size_t)P.ni[i]])) {
                    vi = s.vertex;
                    found = true;
                }
        } else if (!cand.empty()) {
            vi = cand.front().vertex;  // the first match (break)
            found = true;
        }
        if (!found) {
            vi = (uint32_t)O.v.size();
            const V2 uv = has_uv ? P.uv[(size_t)ti] : V2{ 0.f, 0.f };
            const V3 nr = has_n ? P.nrm[(size_t)P.ni[i]] : V3{ 0.f, 0.f, 0.f };
            O.v.push_back(make_vertex(P.pos[(size_t)P.pi[i]], uv, nr));
            cand.push_back(Seen{ has_n ? P.ni[i] : 0, vi });
        }
        O.idx.push_back(vi);
    }
    if (has_uv && has_n) {
        O.type |= RIO_VERTEX_TANGENT;
        for (size_t i = 0; i + 2 < n; i += 3) {  // CCR/ObjParser.cpp:446-469
            rast_ivertex &a = O.v[O.idx[i]], &b = O.v[O.idx[i + 1]], &c = O.v[O.idx[i + 2]];
            const V3 e0{ b.p[0] - a.p[0], b.p[1] - a.p[1], b.p[2] - a.p[2] };
            const V3 e1{ c.p[0] - a.p[0], c.p[1] - a.p[1], c.p[2] - a.p[2] };
            const V2 dx{ b.uv[0] - a.uv[0], c.uv[0] - a.uv[0] };
            const V2 dy{ b.uv[1] - a.uv[1], c.uv[1] - a.uv[1] };
            const float r = 1.f / (dx.x * dy.y - dx.y * dy.x);
            // (edge0 * diffY.y - edge1 * diffY.x) * r
            const V3 t{ (e0.x * dy.y - e1.x * dy.x) * r, (e0.y * dy.y - e1.y * dy.x) * r,
                        (e0.z * dy.y - e1.z * dy.x) * r };
            for (rast_ivertex *v : { &a, &b, &c }) {
                v->tan[0] += t.x;
                v->tan[1] += t.y;
                v->tan[2] += t.z;
            }
        }
        for (rast_ivertex &v : O.v) {  // CCR/ObjParser.cpp:470-473
            const V3 t = normalized(reject(V3{ v.tan[0], v.tan[1], v.tan[2] }, V3{ v.n[0], v.n[1], v.n[2] }));
            v.tan[0] = t.x;
            v.tan[1] = t.y;
            v.tan[2] = t.z;
        }
    }
    return RAST_OK;
}

}  // namespace

extern "C" int rio_last_error(char *buf, size_t n)
{
    if (!buf || !n) return RAST_E_INVALID_ARG;
    std::snprintf(buf, n, "%s", g_err.c_str());
    return RAST_OK;
}

extern "C" void rio_free(void *p) { std::free(p); }

extern "C" int rio_obj_parse(const char *text, size_t len, int invert_y, rast_ivertex **verts, uint32_t *num_verts,
                             uint32_t **indices, uint32_t *num_indices, int *vertex_type)
{
    if (!text || !verts || !num_verts || !indices || !num_indices) return fail("rio_obj_parse: null argument");
    *verts = nullptr;
    *indices = nullptr;
    *num_verts = *num_indices = 0;
    Parsed P;
    int rc = parse(text, len, invert_y != 0, P);
    if (rc != RAST_OK) return rc;
    if (P.pi.empty()) return fail("no faces");
    Out O;
    rc = assign(P, O);
    if (rc != RAST_OK) return rc;
    if (O.v.size() >= 0xFFFFFFFFu || O.idx.size() >= 0xFFFFFFFFu) return fail("mesh too large");
    rast_ivertex *v = (rast_ivertex *)std::malloc(O.v.size() * sizeof(rast_ivertex));
    uint32_t *ix = (uint32_t *)std::malloc(O.idx.size() * sizeof(uint32_t));
    if (!v || !ix) {
        std::free(v);
        std::free(ix);
        g_err = "out of host memory";
        return RAST_E_OUT_OF_MEMORY;
    }
    std::memcpy(v, O.v.data(), O.v.size() * sizeof(rast_ivertex));
    std::memcpy(ix, O.idx.data(), O.idx.size() * sizeof(uint32_t));
    *verts = v;
    *indices = ix;
    *num_verts = (uint32_t)O.v.size();
    *num_indices = (uint32_t)O.idx.size();
    if (vertex_type) *vertex_type = O.type;
    return RAST_OK;
}

extern "C" int rio_obj_load(const char *path, int invert_y, rast_ivertex **verts, uint32_t *num_verts,
                            uint32_t **indices, uint32_t *num_indices, int *vertex_type)
{
    if (!path) return fail("rio_obj_load: null path");
    // OpenFile: the extension after the first '.' must be ".obj" (CCR/ObjParser.cpp:564-584)
    const char *dot = std::strchr(path, '.');
    if (!dot || std::strcmp(dot, ".obj") != 0) return fail("only '.obj' files are supported: %s", path);
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail("cannot open %s", path);
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    std::fclose(f);
    return rio_obj_parse(text.data(), text.size(), invert_y, verts, num_verts, indices, num_indices, vertex_type);
}

extern "C" int rio_write_bmp(const char *path, const uint32_t *rgba8, uint32_t width, uint32_t height)
{
    if (!path || !rgba8 || !width || !height || width > 65535 || height > 65535)
        return fail("rio_write_bmp: bad arguments");
    const uint64_t row = (uint64_t)width * 4, data = row * height;
    const uint32_t off = 14 + 40;
    unsigned char hdr[54] = {};
    auto put16 = [&](int at, uint32_t v) { hdr[at] = (unsigned char)v; hdr[at + 1] = (unsigned char)(v >> 8); };
    auto put32 = [&](int at, uint32_t v) { put16(at, v & 0xFFFF); put16(at + 2, v >> 16); };
    hdr[0] = 'B';
    hdr[1] = 'M';
    put32(2, (uint32_t)(off + data));
    put32(10, off);
    put32(14, 40);           // BITMAPINFOHEADER
    put32(18, width);
    put32(22, height);       // positive: bottom-up rows
    put16(26, 1);            // planes
    put16(28, 32);           // bits per pixel
    put32(30, 0);            // BI_RGB
    put32(34, (uint32_t)data);
    put32(38, 2835);         // 72 dpi
    put32(42, 2835);
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail("cannot create %s", path);
    bool ok = std::fwrite(hdr, 1, sizeof hdr, f) == sizeof hdr;
    std::vector<unsigned char> line((size_t)row);
    for (uint32_t y = 0; ok && y < height; ++y) {
        const uint32_t *src = rgba8 + (size_t)(height - 1 - y) * width;
        for (uint32_t x = 0; x < width; ++x) {
            const uint32_t t = src[x];
            line[(size_t)x * 4 + 0] = (unsigned char)(t >> 16);  // B
            line[(size_t)x * 4 + 1] = (unsigned char)(t >> 8);   // G
            line[(size_t)x * 4 + 2] = (unsigned char)t;          // R
            line[(size_t)x * 4 + 3] = (unsigned char)(t >> 24);  // A
        }
        ok = std::fwrite(line.data(), 1, line.size(), f) == line.size();
    }
    if (std::fclose(f) != 0) ok = false;
    return ok ? RAST_OK : fail("write to %s failed", path);
}

// ------------------------------------------------------------------- PNG
//
// The reference loads its textures with SDL_image's IMG_Load
// (CCR/CUDATexture.cpp:24-45), which for its PNG files yields 32-bit RGBA
// texels (bytes R, G, B, A in memory: the texel format used here).  This
// decoder is written from the PNG specification (ISO/IEC 15948): signature,
// chunks with their CRC-32, IHDR, PLTE / tRNS, the concatenated IDAT stream
// inflated with zlib, and the five per-row filters (None, Sub, Up, Average,
// Paeth) undone over the bytes of a row.  Bit depth 8, colour types 0
// (grey), 2 (RGB), 3 (palette), 4 (grey + alpha) and 6 (RGBA), no
// interlace; anything else is rejected with an error.

namespace {

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

void put_be32(std::vector<uint8_t> &o, uint32_t v)
{
    o.push_back((uint8_t)(v >> 24));
    o.push_back((uint8_t)(v >> 16));
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

const uint8_t kPngSig[8] = { 137, 80, 78, 71, 13, 10, 26, 10 };

// Paeth predictor (ISO/IEC 15948, 9.4): the neighbour nearest to a + b - c.
inline uint8_t paeth(int a, int b, int c)
{
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    return (uint8_t)(pb <= pc ? b : c);
}

// Undoes filter `type` of one row in place; prev is the previous row after
// unfiltering (zeros above the first row), bpp the bytes per pixel.
bool unfilter_row(int type, uint8_t *row, const uint8_t *prev, size_t n, size_t bpp)
{
    switch (type) {
    case 0: return true;
    case 1:
        for (size_t i = bpp; i < n; ++i) row[i] = (uint8_t)(row[i] + row[i - bpp]);
        return true;
    case 2:
        for (size_t i = 0; i < n; ++i) row[i] = (uint8_t)(row[i] + prev[i]);
        return true;
    case 3:
        for (size_t i = 0; i < n; ++i) {
            const int a = i >= bpp ? row[i - bpp] : 0;
            row[i] = (uint8_t)(row[i] + ((a + prev[i]) >> 1));
        }
        return true;
    case 4:
        for (size_t i = 0; i < n; ++i) {
            const int a = i >= bpp ? row[i - bpp] : 0, c = i >= bpp ? prev[i - bpp] : 0;
            row[i] = (uint8_t)(row[i] + paeth(a, prev[i], c));
        }
        return true;
    default: return false;
    }
}

// Applies filter `type` to one row (the inverse of unfilter_row): out[i] =
// raw[i] - predictor, predictors from the unfiltered bytes.
void filter_row(int type, const uint8_t *raw, const uint8_t *prev, size_t n, size_t bpp, uint8_t *out)
{
    for (size_t i = 0; i < n; ++i) {
        const int a = i >= bpp ? raw[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
        int pred = 0;
        switch (type) {
        case 1: pred = a; break;
        case 2: pred = b; break;
        case 3: pred = (a + b) >> 1; break;
        case 4: pred = paeth(a, b, c); break;
        default: pred = 0; break;
        }
        out[i] = (uint8_t)(raw[i] - pred);
    }
}

}  // namespace

extern "C" int rio_png_decode(const uint8_t *data, size_t len, uint32_t **texels, uint32_t *width, uint32_t *height)
{
    if (!data || !texels || !width || !height) return fail("rio_png_decode: null argument");
    *texels = nullptr;
    *width = *height = 0;
    if (len < 8 || std::memcmp(data, kPngSig, 8) != 0) return fail("not a PNG file (signature)");
    size_t at = 8;
    uint32_t w = 0, h = 0;
    int ctype = -1, seen_ihdr = 0, seen_end = 0;
    uint8_t pal[256][4];
    uint32_t npal = 0;
    for (int i = 0; i < 256; ++i) pal[i][0] = pal[i][1] = pal[i][2] = 0, pal[i][3] = 255;
    int trns_grey = -1, trns_rgb[3] = { -1, -1, -1 };
    std::vector<uint8_t> z;
    while (!seen_end) {
        if (len - at < 12) return fail("truncated chunk at byte %zu", at);
        const uint32_t n = be32(data + at);
        const uint8_t *type = data + at + 4, *body = data + at + 8;
        if (n > 0x7FFFFFFFu || len - at - 12 < n) return fail("chunk length %u past the end of the file", n);
        const uint32_t crc = be32(body + n);
        if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, n + 4) != crc)
            return fail("CRC mismatch in chunk %.4s", (const char *)type);
        const bool critical = !(type[0] & 0x20);
        if (std::memcmp(type, "IHDR", 4) == 0) {
            if (seen_ihdr || n != 13) return fail("bad IHDR");
            w = be32(body);
            h = be32(body + 4);
            const int depth = body[8];
            ctype = body[9];
            if (w == 0 || h == 0 || w > 65535 || h > 65535) return fail("image size %ux%u not supported", w, h);
            if (depth != 8) return fail("bit depth %d not supported (8 only)", depth);
            if (ctype != 0 && ctype != 2 && ctype != 3 && ctype != 4 && ctype != 6) return fail("colour type %d", ctype);
            if (body[10] != 0 || body[11] != 0) return fail("unknown compression or filter method");
            if (body[12] != 0) return fail("interlaced PNG not supported");
            seen_ihdr = 1;
        } else if (!seen_ihdr) {
            return fail("first chunk is not IHDR");
        } else if (std::memcmp(type, "PLTE", 4) == 0) {
            if (n % 3 || n / 3 > 256 || n == 0) return fail("bad PLTE");
            npal = n / 3;
            for (uint32_t i = 0; i < npal; ++i)
                for (int c = 0; c < 3; ++c) pal[i][c] = body[3 * i + c];
        } else if (std::memcmp(type, "tRNS", 4) == 0) {
            if (ctype == 3) {
                if (n > npal) return fail("tRNS longer than the palette");
                for (uint32_t i = 0; i < n; ++i) pal[i][3] = body[i];
            } else if (ctype == 0 && n == 2) {
                trns_grey = (int)(body[0] << 8 | body[1]);
            } else if (ctype == 2 && n == 6) {
                for (int c = 0; c < 3; ++c) trns_rgb[c] = (int)(body[2 * c] << 8 | body[2 * c + 1]);
            } else {
                return fail("bad tRNS");
            }
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            z.insert(z.end(), body, body + n);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            seen_end = 1;
        } else if (critical) {
            return fail("unknown critical chunk %.4s", (const char *)type);
        }
        at += 12 + (size_t)n;
    }
    if (z.empty()) return fail("no IDAT");
    if (ctype == 3 && npal == 0) return fail("palette image without PLTE");
    const size_t bpp = ctype == 0 || ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 2 ? 3 : 4;
    const size_t stride = (size_t)w * bpp, raw_len = (stride + 1) * h;
    std::vector<uint8_t> raw(raw_len + 1);
    uLongf got = (uLongf)raw.size();
    const int zr = uncompress(raw.data(), &got, z.data(), (uLong)z.size());
    if (zr != Z_OK) return fail("zlib inflate failed (%d)", zr);
    if (got != raw_len) return fail("image data is %lu bytes, expected %zu", (unsigned long)got, raw_len);
    uint32_t *out = (uint32_t *)std::malloc((size_t)w * h * 4);
    if (!out) {
        g_err = "out of host memory";
        return RAST_E_OUT_OF_MEMORY;
    }
    std::vector<uint8_t> zero(stride, 0);
    const uint8_t *prev = zero.data();
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t *row = raw.data() + (size_t)y * (stride + 1);
        if (!unfilter_row(row[0], row + 1, prev, stride, bpp)) {
            std::free(out);
            return fail("row %u: unknown filter type %d", y, row[0]);
        }
        const uint8_t *px = row + 1;
        uint32_t *dst = out + (size_t)y * w;
        for (uint32_t x = 0; x < w; ++x) {
            uint32_t r, g, b, a = 255;
            switch (ctype) {
            case 0:
                r = g = b = px[x];
                if (trns_grey == (int)px[x]) a = 0;
                break;
            case 2:
                r = px[3 * x], g = px[3 * x + 1], b = px[3 * x + 2];
                if (trns_rgb[0] == (int)r && trns_rgb[1] == (int)g && trns_rgb[2] == (int)b) a = 0;
                break;
            case 3:
                if (px[x] >= npal) {
                    std::free(out);
                    return fail("palette index %u out of range", px[x]);
                }
                r = pal[px[x]][0], g = pal[px[x]][1], b = pal[px[x]][2], a = pal[px[x]][3];
                break;
            case 4: r = g = b = px[2 * x], a = px[2 * x + 1]; break;
            default: r = px[4 * x], g = px[4 * x + 1], b = px[4 * x + 2], a = px[4 * x + 3]; break;
            }
            dst[x] = r | g << 8 | b << 16 | a << 24;
        }
        prev = row + 1;
    }
    *texels = out;
    *width = w;
    *height = h;
    return RAST_OK;
}

extern "C" int rio_png_load(const char *path, uint32_t **texels, uint32_t *width, uint32_t *height)
{
    if (!path) return fail("rio_png_load: null path");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail("cannot open %s", path);
    std::vector<uint8_t> bytes;
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) bytes.insert(bytes.end(), buf, buf + got);
    std::fclose(f);
    return rio_png_decode(bytes.data(), bytes.size(), texels, width, height);
}

extern "C" int rio_png_encode(const uint32_t *rgba8, uint32_t width, uint32_t height, int filter, uint8_t **out,
                              size_t *out_len)
{
    if (!rgba8 || !out || !out_len || !width || !height || width > 65535 || height > 65535 || filter < -1 ||
        filter > 4)
        return fail("rio_png_encode: bad arguments");
    *out = nullptr;
    *out_len = 0;
    const size_t stride = (size_t)width * 4;
    std::vector<uint8_t> raw((stride + 1) * height), cur(stride), prev(stride, 0);
    for (uint32_t y = 0; y < height; ++y) {
        for (uint32_t x = 0; x < width; ++x) {
            const uint32_t t = rgba8[(size_t)y * width + x];
            for (int c = 0; c < 4; ++c) cur[4 * (size_t)x + c] = (uint8_t)(t >> (8 * c));
        }
        const int type = filter >= 0 ? filter : (int)(y % 5);  // -1: every filter in turn
        uint8_t *row = raw.data() + (size_t)y * (stride + 1);
        row[0] = (uint8_t)type;
        filter_row(type, cur.data(), prev.data(), stride, 4, row + 1);
        prev.swap(cur);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return fail("zlib deflate failed");
    std::vector<uint8_t> o(kPngSig, kPngSig + 8);
    auto chunk = [&](const char *type, const uint8_t *body, size_t n) {
        put_be32(o, (uint32_t)n);
        const size_t start = o.size();
        o.insert(o.end(), type, type + 4);
        o.insert(o.end(), body, body + n);
        put_be32(o, (uint32_t)crc32(crc32(0L, Z_NULL, 0), o.data() + start, (uInt)(n + 4)));
    };
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, width);
    put_be32(ihdr, height);
    const uint8_t tail[5] = { 8, 6, 0, 0, 0 };  // 8-bit RGBA, deflate, adaptive filtering, no interlace
    ihdr.insert(ihdr.end(), tail, tail + 5);
    chunk("IHDR", ihdr.data(), ihdr.size());
    chunk("IDAT", z.data(), zlen);
    chunk("IEND", nullptr, 0);
    uint8_t *p = (uint8_t *)std::malloc(o.size());
    if (!p) {
        g_err = "out of host memory";
        return RAST_E_OUT_OF_MEMORY;
    }
    std::memcpy(p, o.data(), o.size());
    *out = p;
    *out_len = o.size();
    return RAST_OK;
}
