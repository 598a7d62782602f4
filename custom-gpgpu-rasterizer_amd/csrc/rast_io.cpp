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
