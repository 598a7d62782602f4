// scenes.cpp -- deterministic synthetic inputs (meshes, textures, cameras).
//
// Host-only.  Uses IEEE basic operations, sqrt and floor plus rsc_sin /
// rsc_cos below; built with -ffp-contract=off so the outputs are the same
// bytes on every x86-64 host.  Vertex layout is the reference IVertex
// (CCR/Vertex.h:113-145): p[3] uv[2] n[3] tan[3] c[3] = 14 floats.
#include "rast_scene.h"

#include <cmath>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

namespace {

constexpr int kVF = 14;  // floats per vertex
constexpr double kPi = 3.14159265358979323846;

uint32_t hash32(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352dU;
    x ^= x >> 15; x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

double lattice(int32_t ix, int32_t iy, uint32_t seed)
{
    uint32_t h = hash32((uint32_t)ix * 0x8da6b343U ^ (uint32_t)iy * 0xd8163841U ^ hash32(seed + 0x9e3779b9U));
    return (double)(h >> 8) * (1.0 / 16777216.0);
}

double value_noise(double x, double y, uint32_t seed)
{
    double fx0 = std::floor(x), fy0 = std::floor(y);
    int32_t ix = (int32_t)fx0, iy = (int32_t)fy0;
    double tx = x - fx0, ty = y - fy0;
    tx = tx * tx * (3.0 - 2.0 * tx);
    ty = ty * ty * (3.0 - 2.0 * ty);
    double a = lattice(ix, iy, seed), b = lattice(ix + 1, iy, seed);
    double c = lattice(ix, iy + 1, seed), d = lattice(ix + 1, iy + 1, seed);
    double ab = a + (b - a) * tx, cd = c + (d - c) * tx;
    return ab + (cd - ab) * ty;
}

double fbm(double x, double y, uint32_t seed, uint32_t octaves)
{
    double sum = 0.0, amp = 0.5, norm = 0.0, f = 1.0;
    for (uint32_t o = 0; o < octaves; ++o) {
        sum += amp * value_noise(x * f, y * f, seed + o * 1013U);
        norm += amp;
        amp *= 0.5;
        f *= 2.0;
    }
    return sum / norm;
}

struct V3 { double x, y, z; };
V3 sub(V3 a, V3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
V3 add(V3 a, V3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
V3 mul(V3 a, double s) { return { a.x * s, a.y * s, a.z * s }; }
V3 cross(V3 a, V3 b) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 norm(V3 a)
{
    double m = std::sqrt(dot(a, a));
    return m > 0 ? mul(a, 1.0 / m) : V3{ 0, 0, 0 };
}

void put_vertex(float *dst, V3 p, double u, double v, V3 n, V3 t, V3 c)
{
    const double f[kVF] = { p.x, p.y, p.z, u, v, n.x, n.y, n.z, t.x, t.y, t.z, c.x, c.y, c.z };
    for (int i = 0; i < kVF; ++i) dst[i] = (float)f[i];
}

V3 tangent_for(V3 n)
{
    V3 t = cross(V3{ 0, 1, 0 }, n);
    if (dot(t, t) < 1e-12) t = V3{ 1, 0, 0 };
    return norm(t);
}

uint8_t to_u8(double v)
{
    if (!(v > 0.0)) return 0;
    if (v >= 1.0) return 255;
    return (uint8_t)(v * 255.0);
}

uint32_t rgba(double r, double g, double b)
{
    return (uint32_t)to_u8(r) | ((uint32_t)to_u8(g) << 8) | ((uint32_t)to_u8(b) << 16) | 0xFF000000u;
}

}  // namespace

extern "C" {

double rsc_sin(double x)
{
    // Cody-Waite reduction by pi/2, then Taylor polynomials on [-pi/4, pi/4].
    const double pio2_hi = 1.57079632673412561417e+00;
    const double pio2_lo = 6.07710050650619224932e-11;
    double k = std::floor(x * (2.0 / kPi) + 0.5);
    double r = (x - k * pio2_hi) - k * pio2_lo;
    double r2 = r * r;
    double s = r * (1.0 + r2 * (-1.0 / 6 + r2 * (1.0 / 120 + r2 * (-1.0 / 5040 + r2 * (1.0 / 362880 +
               r2 * (-1.0 / 39916800 + r2 * (1.0 / 6227020800.0 + r2 * (-1.0 / 1307674368000.0))))))));
    double c = 1.0 + r2 * (-0.5 + r2 * (1.0 / 24 + r2 * (-1.0 / 720 + r2 * (1.0 / 40320 +
               r2 * (-1.0 / 3628800 + r2 * (1.0 / 479001600.0 + r2 * (-1.0 / 87178291200.0)))))));
    long q = (long)k & 3;
    if (q < 0) q += 4;
    switch (q) {
    case 0: return s;
    case 1: return c;
    case 2: return -s;
    default: return -c;
    }
}

double rsc_cos(double x) { return rsc_sin(x + kPi / 2); }

float rsc_uniform(uint32_t seed, uint32_t index, float lo, float hi)
{
    uint32_t h = hash32(hash32(seed) ^ (index * 0x9e3779b9U + 0x7f4a7c15U));
    double u = (double)(h >> 8) * (1.0 / 16777216.0);
    return (float)(lo + (hi - lo) * u);
}

int rsc_uv_sphere(uint32_t slices, uint32_t stacks, float radius, float *verts, uint32_t *num_verts,
                  uint32_t *indices, uint32_t *num_indices)
{
    if (slices < 3 || stacks < 2) return -1;
    const uint32_t nv = (stacks + 1) * (slices + 1);
    const uint32_t ntri = slices * (2 * stacks - 1);
    if (num_verts) *num_verts = nv;
    if (num_indices) *num_indices = 3 * ntri;
    if (!verts || !indices) return 0;
    for (uint32_t i = 0; i <= stacks; ++i) {
        double th = kPi * i / stacks;
        double st = rsc_sin(th), ct = rsc_cos(th);
        if (i == 0) { st = 0; ct = 1; }
        if (i == stacks) { st = 0; ct = -1; }
        for (uint32_t j = 0; j <= slices; ++j) {
            double ph = 2 * kPi * j / slices;
            double sp = rsc_sin(ph), cp = rsc_cos(ph);
            // phi = 0 faces -z so the texture seam (u = 0 / 1) is at the back
            V3 n{ -st * sp, ct, -st * cp };
            V3 t{ -cp, 0, sp };
            put_vertex(verts + (size_t)kVF * (i * (slices + 1) + j), mul(n, radius), (double)j / slices,
                       (double)i / stacks, n, t, V3{ 1, 1, 1 });
        }
    }
    uint32_t *ix = indices;
    for (uint32_t i = 0; i < stacks; ++i)
        for (uint32_t j = 0; j < slices; ++j) {
            uint32_t a = i * (slices + 1) + j, b = a + 1, c = a + slices + 1, d = c + 1;
            *ix++ = a; *ix++ = c; *ix++ = d;  // counter-clockwise seen from outside
            if (i != 0) { *ix++ = a; *ix++ = d; *ix++ = b; }
        }
    return 0;
}

int rsc_icosphere(uint32_t subdiv, float radius, float *verts, uint32_t *num_verts, uint32_t *indices,
                  uint32_t *num_indices)
{
    if (subdiv > 9) return -1;
    uint32_t p4 = 1u << (2 * subdiv);
    if (num_verts) *num_verts = 10 * p4 + 2;
    if (num_indices) *num_indices = 3 * 20 * p4;
    if (!verts || !indices) return 0;
    const double g = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<V3> P = { { -1, g, 0 }, { 1, g, 0 }, { -1, -g, 0 }, { 1, -g, 0 },
                          { 0, -1, g }, { 0, 1, g }, { 0, -1, -g }, { 0, 1, -g },
                          { g, 0, -1 }, { g, 0, 1 }, { -g, 0, -1 }, { -g, 0, 1 } };
    for (auto &p : P) p = norm(p);
    std::vector<uint32_t> F = { 0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                                11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                                3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1 };
    for (size_t f = 0; f < F.size(); f += 3) {  // make every face counter-clockwise from outside
        V3 a = P[F[f]], b = P[F[f + 1]], c = P[F[f + 2]];
        if (dot(cross(sub(b, a), sub(c, a)), add(add(a, b), c)) < 0) std::swap(F[f + 1], F[f + 2]);
    }
    for (uint32_t s = 0; s < subdiv; ++s) {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            auto key = std::make_pair(a < b ? a : b, a < b ? b : a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            uint32_t id = (uint32_t)P.size();
            P.push_back(norm(mul(add(P[a], P[b]), 0.5)));
            mid.emplace(key, id);
            return id;
        };
        std::vector<uint32_t> NF;
        NF.reserve(F.size() * 4);
        for (size_t f = 0; f < F.size(); f += 3) {
            uint32_t a = F[f], b = F[f + 1], c = F[f + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            const uint32_t tris[12] = { a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca };
            NF.insert(NF.end(), tris, tris + 12);
        }
        F.swap(NF);
    }
    for (size_t i = 0; i < P.size(); ++i) {
        V3 n = P[i];
        put_vertex(verts + kVF * i, mul(n, radius), 0.5 + 0.5 * n.x, 0.5 - 0.5 * n.y, n, tangent_for(n),
                   V3{ 1, 1, 1 });
    }
    std::memcpy(indices, F.data(), F.size() * sizeof(uint32_t));
    return 0;
}

int rsc_terrain(uint32_t n, float size, float amplitude, uint32_t octaves, uint32_t seed, float uv_tiling,
                float *verts, uint32_t *num_verts, uint32_t *indices, uint32_t *num_indices)
{
    if (n < 1 || n > 16383) return -1;
    const uint32_t m = n + 1;
    if (num_verts) *num_verts = m * m;
    if (num_indices) *num_indices = 6 * n * n;
    if (!verts || !indices) return 0;
    std::vector<double> h((size_t)m * m);
    const double freq = 8.0;
    for (uint32_t j = 0; j < m; ++j)
        for (uint32_t i = 0; i < m; ++i)
            h[(size_t)j * m + i] = amplitude * fbm(freq * i / n, freq * j / n, seed, octaves);
    const double dx = (double)size / n;
    for (uint32_t j = 0; j < m; ++j)
        for (uint32_t i = 0; i < m; ++i) {
            uint32_t i0 = i ? i - 1 : i, i1 = i < n ? i + 1 : i;
            uint32_t j0 = j ? j - 1 : j, j1 = j < n ? j + 1 : j;
            double dhdx = (h[(size_t)j * m + i1] - h[(size_t)j * m + i0]) / (dx * (i1 - i0));
            double dhdz = (h[(size_t)j1 * m + i] - h[(size_t)j0 * m + i]) / (dx * (j1 - j0));
            V3 nn = norm(V3{ -dhdx, 1.0, -dhdz });
            V3 t = norm(V3{ 1.0, dhdx, 0.0 });
            V3 p{ -0.5 * size + dx * i, h[(size_t)j * m + i], -0.5 * size + dx * j };
            // texel centres stay inside (0, uv_tiling): the reference sampler does not wrap
            put_vertex(verts + (size_t)kVF * ((size_t)j * m + i), p, uv_tiling * (i + 0.5) / (n + 1),
                       uv_tiling * (j + 0.5) / (n + 1), nn, t, V3{ 1, 1, 1 });
        }
    uint32_t *ix = indices;
    for (uint32_t j = 0; j < n; ++j)
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t a = j * m + i, b = a + 1, c = a + m, d = c + 1;
            *ix++ = a; *ix++ = c; *ix++ = b;  // normal +y by the right-hand rule
            *ix++ = b; *ix++ = c; *ix++ = d;
        }
    return 0;
}

int rsc_texture_checker(uint32_t w, uint32_t h, uint32_t square, uint32_t c0, uint32_t c1, uint32_t *texels)
{
    if (!w || !h || !square || !texels) return -1;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) texels[(size_t)y * w + x] = (((x / square) + (y / square)) & 1u) ? c1 : c0;
    return 0;
}

int rsc_texture_noise(uint32_t w, uint32_t h, uint32_t seed, int kind, uint32_t *texels)
{
    if (!w || !h || !texels || kind < 0 || kind > 3) return -1;
    const double cells = 16.0;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            double u = cells * x / w, v = cells * y / h;
            uint32_t out;
            if (kind == 0) {
                double a = fbm(u, v, seed, 5), b = fbm(u, v, seed + 77, 5), c = fbm(u, v, seed + 191, 5);
                out = rgba(0.25 + 0.7 * a, 0.2 + 0.6 * b, 0.15 + 0.5 * c);
            } else if (kind == 1) {
                double e = cells / w;
                double hx = fbm(u + e, v, seed, 4) - fbm(u - e, v, seed, 4);
                double hy = fbm(u, v + e, seed, 4) - fbm(u, v - e, seed, 4);
                V3 nn = norm(V3{ -hx * 40.0, -hy * 40.0, 1.0 });
                out = rgba(0.5 + 0.5 * nn.x, 0.5 + 0.5 * nn.y, 0.5 + 0.5 * nn.z);
            } else {
                double a = fbm(u, v, seed + (uint32_t)kind * 31U, 4);
                double s = kind == 2 ? 0.2 + 0.6 * a : 0.3 + 0.7 * a;
                out = rgba(s, s, s);
            }
            texels[(size_t)y * w + x] = out;
        }
    return 0;
}

void rsc_world_matrix(const float t[3], const float e[3], float s, float out[16])
{
    // R = Rz * Ry * Rx (CCR/Matrix3.h:356-362 order), M = T * R * S.
    double cx = rsc_cos(e[0]), sx = rsc_sin(e[0]);
    double cy = rsc_cos(e[1]), sy = rsc_sin(e[1]);
    double cz = rsc_cos(e[2]), sz = rsc_sin(e[2]);
    double Rx[3][3] = { { 1, 0, 0 }, { 0, cx, -sx }, { 0, sx, cx } };
    double Ry[3][3] = { { cy, 0, sy }, { 0, 1, 0 }, { -sy, 0, cy } };
    double Rz[3][3] = { { cz, -sz, 0 }, { sz, cz, 0 }, { 0, 0, 1 } };
    double A[3][3], R[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) A[r][c] = Rz[r][0] * Ry[0][c] + Rz[r][1] * Ry[1][c] + Rz[r][2] * Ry[2][c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) R[r][c] = A[r][0] * Rx[0][c] + A[r][1] * Rx[1][c] + A[r][2] * Rx[2][c];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double v;
            if (r < 3 && c < 3) v = R[r][c] * s;
            else if (r < 3 && c == 3) v = t[r];
            else v = (c == 3) ? 1.0 : 0.0;
            out[c * 4 + r] = (float)v;  // data[col][row]
        }
}

int rsc_camera_path(uint32_t seed, uint32_t frames, float terrain_size, float height, float pitch_deg,
                    float *out)
{
    if (!out || !frames) return -1;
    double ph[6], fr[6];
    for (int k = 0; k < 6; ++k) {
        ph[k] = rsc_uniform(seed, 10 + k, 0.0f, (float)(2 * kPi));
        fr[k] = rsc_uniform(seed, 20 + k, 0.5f, 1.5f);
    }
    // The camera stays outside the terrain's +z edge with |yaw| <= 0.12 rad, so
    // every terrain vertex is in front of the eye plane (the reference does not
    // clip at the near plane, CCR/CUDARenderer.cu:702-796).
    const double half = 0.5 * terrain_size;
    for (uint32_t f = 0; f < frames; ++f) {
        double t = 2 * kPi * f / frames;
        double x = 0.3 * half * rsc_sin(fr[0] * t + ph[0]);
        double z = half + 0.25 * half * (1.2 + 0.5 * rsc_sin(fr[1] * t + ph[1]));
        double y = height * (1.0 + 0.1 * rsc_sin(fr[2] * t + ph[2]));
        double yaw = 0.12 * rsc_sin(fr[3] * t + ph[3]);
        double pitch = -(pitch_deg + 2.0 * rsc_sin(fr[4] * t + ph[4])) * kPi / 180.0;
        V3 d{ rsc_sin(yaw) * rsc_cos(pitch), rsc_sin(pitch), -rsc_cos(yaw) * rsc_cos(pitch) };
        float *o = out + 6 * (size_t)f;
        o[0] = (float)x; o[1] = (float)y; o[2] = (float)z;
        o[3] = (float)-d.x; o[4] = (float)-d.y; o[5] = (float)-d.z;
    }
    return 0;
}

int rsc_backdrop(const float cam_pos[3], const float cam_forward[3], float fov_deg, float aspect, float d0,
                 float step, float *verts, uint32_t *indices)
{
    if (!verts || !indices) return -1;
    V3 fwd = norm(V3{ cam_forward[0], cam_forward[1], cam_forward[2] });
    V3 right = norm(cross(V3{ 0, 1, 0 }, fwd));
    V3 up = cross(fwd, right);
    V3 view = mul(fwd, -1.0);
    V3 cp{ cam_pos[0], cam_pos[1], cam_pos[2] };
    const double tanh = rsc_sin(fov_deg * kPi / 360.0) / rsc_cos(fov_deg * kPi / 360.0);
    for (int k = 0; k < 8; ++k) {
        double D = d0 + step * k;
        double hh = 1.2 * (D + 1.0) * tanh, hw = hh * aspect;
        const double xy[3][2] = { { -3 * hw, -1.1 * hh }, { 3 * hw, -1.1 * hh }, { 0, 3 * hh } };
        V3 centre = add(cp, mul(view, D));
        for (int v = 0; v < 3; ++v) {
            V3 p = add(add(centre, mul(right, xy[v][0])), mul(up, xy[v][1]));
            put_vertex(verts + kVF * (3 * k + v), p, 0.5 + xy[v][0] / (6 * hw), 0.5 - xy[v][1] / (6 * hh), fwd,
                       right, V3{ 0.6, 0.7, 0.9 });
            indices[3 * k + v] = (uint32_t)(3 * k + v);
        }
    }
    return 0;
}

}  // extern "C"
