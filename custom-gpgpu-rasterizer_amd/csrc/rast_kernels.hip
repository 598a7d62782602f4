// rast_kernels.hip -- the per-frame rasterization kernels for gfx950.
//
// Built with -ffp-contract=off and no fast-math: every float expression below
// is evaluated as written (one IEEE binary32 operation each, left to right),
// division and sqrt correctly rounded, so visibility and depth are bitwise the
// CPU restatement of the reference path.  Reference files are cited relative
// to kdedeurw/Custom-GPGPU-Rasterizer (CCR = CustomCUDARasterizer/).
//
// Pipeline (one launch per stage, all meshes at once):
//   k_vertex   per vertex: clip/NDC position + shading attributes
//   k_setup    per triangle: assembly, cull, screen setup record, then either
//              tile counts (exact two-pass) or direct queue fill (single pass)
//   k_scan     per-tile exclusive scan of counts / capacities
//   k_fill     exact second pass: ordered-by-nothing queue fill
//   k_fallback single pass: refill tiles whose capacity overflowed
//   k_raster   one workgroup per tile, depth resolved in LDS
//   k_shade    per pixel, deferred from the visibility buffer
#include <hip/hip_runtime.h>

#include "rast_launch.h"

namespace rast {

namespace {

// ------------------------------------------------------------ wave helpers

__device__ __forceinline__ uint32_t lane_id()
{
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of m below this lane.
__device__ __forceinline__ uint32_t popc_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Order-free append of one element per lane with pred set: one atomic per
// wave (ballot + prefix count).  Returns the lane's slot (garbage if !pred).
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred)
{
    const uint64_t m = __ballot(pred);
    if (m == 0) return 0;
    const uint32_t first = (uint32_t)__builtin_ctzll(m);
    uint32_t base = 0;
    if (lane_id() == first) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __builtin_amdgcn_readlane(base, first);
    return base + popc_below(m);
}

// Per-lane atomic increment of counters[t] aggregated over the lanes of the
// wave that hit the same t ("peeling": one atomic per distinct t).  Must be
// called by the active lanes only; returns the lane's previous-value slot.
__device__ __forceinline__ uint32_t agg_increment(uint32_t *counters, uint32_t t)
{
    uint32_t slot = 0;
    bool done = false;
    while (!done) {
        const uint32_t lt = __builtin_amdgcn_readfirstlane(t);
        const uint64_t m = __ballot(t == lt);
        if (t == lt) {
            const uint32_t rank = popc_below(m);
            uint32_t b = 0;
            if (rank == 0) b = atomicAdd(counters + lt, (uint32_t)__popcll(m));
            b = __builtin_amdgcn_readfirstlane(b);
            slot = b + rank;
            done = true;
        }
    }
    return slot;
}

// ------------------------------------------------- reference arithmetic

__device__ __forceinline__ float m4(const float *m, int r, int c) { return m[c * 4 + r]; }

// Normalize (CCR/Vector.h:41-55): m = sqrt(Dot(v,v)); v *= 1/m; zero stays zero.
__device__ __forceinline__ void normalize3(float &x, float &y, float &z)
{
    const float m = sqrtf(x * x + y * y + z * z);
    if (m == 0.f) { x = 0.f; y = 0.f; z = 0.f; return; }
    const float inv = 1.f / m;
    x *= inv; y *= inv; z *= inv;
}

__device__ __forceinline__ bool all_outside(float a, float b, float c, float lo, float hi)
{
    return (a < lo && b < lo && c < lo) || (a > hi && b > hi && c > hi);
}

// Bbox bound of D4: clamp in float to [-2, hi], truncate toward zero.
__device__ __forceinline__ int clamp_trunc(float v, float hi)
{
    float c = v < -2.f ? -2.f : v;
    c = c > hi ? hi : c;
    return (int)c;
}

constexpr uint64_t kClearKey = ((uint64_t)0x3F800000u << 32) | 0xFFFFFFFFull;  // depth 1.0, no triangle

// Depth-test key: min z first, then max ID (low word = ~(id + 1)).
__device__ __forceinline__ uint64_t make_key(float z, uint32_t id)
{
    const uint32_t zb = __float_as_uint(z) & 0x7FFFFFFFu;  // z in [0,1]; -0.0 -> +0.0
    return ((uint64_t)zb << 32) | (uint32_t)~(id + 1u);
}

// Raster-ready triangle in LDS: screen vertices, the three edge vectors of
// CCR/CUDAROPs.cu:131-133, depths, |area| and the tile-clipped pixel range.
struct TriLds {
    float sx0, sy0, sx1, sy1, sx2, sy2;
    float ax, ay, bx, by, cx, cy;
    float z0, z1, z2, area;
    uint32_t xr, yr;  // x0 | x1 << 16, y0 | y1 << 16 (absolute pixels)
    uint32_t id;
    uint32_t pad;
};

// Fragment test of RasterizePixelAtomic (CCR/CUDAROPs.cu:329-345) with the D2
// window: edge functions (CCR/CUDAROPs.cu:120-145), weights /= |area|,
// z = w0 z0 + w1 z1 + w2 z2, accept 0 <= z <= 1.
__device__ __forceinline__ bool fragment(const TriLds &t, float px, float py, float &z)
{
    const float w2 = (px - t.sx0) * t.ay - (py - t.sy0) * t.ax;
    const float w0 = (px - t.sx1) * t.by - (py - t.sy1) * t.bx;
    const float w1 = (px - t.sx2) * t.cy - (py - t.sy2) * t.cx;
    if (!(w0 >= 0.f && w1 >= 0.f && w2 >= 0.f)) return false;
    const float a0 = w0 / t.area, a1 = w1 / t.area, a2 = w2 / t.area;
    const float zi = (a0 * t.z0) + (a1 * t.z1) + (a2 * t.z2);
    if (!(zi >= 0.f && zi <= 1.f)) return false;
    z = zi;
    return true;
}

__device__ __forceinline__ void load_tri(const uint32_t *__restrict__ rec, uint32_t id, TriLds &t)
{
    const uint4 *r = reinterpret_cast<const uint4 *>(rec + (size_t)id * kTriRecWords);
    const uint4 q0 = r[0], q1 = r[1], q2 = r[2];
    t.sx0 = __uint_as_float(q0.x); t.sy0 = __uint_as_float(q0.y);
    t.sx1 = __uint_as_float(q0.z); t.sy1 = __uint_as_float(q0.w);
    t.sx2 = __uint_as_float(q1.x); t.sy2 = __uint_as_float(q1.y);
    t.z0 = __uint_as_float(q1.z); t.z1 = __uint_as_float(q1.w);
    t.z2 = __uint_as_float(q2.x); t.area = __uint_as_float(q2.y);
    t.xr = q2.z;  // bbox lo (xmin | ymin << 16)
    t.yr = q2.w;  // bbox hi (xmax | ymax << 16)
    t.ax = t.sx1 - t.sx0; t.ay = t.sy1 - t.sy0;
    t.bx = t.sx2 - t.sx1; t.by = t.sy2 - t.sy1;
    t.cx = t.sx0 - t.sx2; t.cy = t.sy0 - t.sy2;
    t.id = id;
    t.pad = 0;
}

// ---------------------------------------------------------------- kernels

__global__ void k_frame_init(uint32_t *tile_count, uint32_t *tile_aux, uint32_t n, FrameCounters *ctr,
                             uint32_t *density)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        tile_count[i] = 0;
        if (tile_aux) tile_aux[i] = 0xFFFFFFFFu;
        if (density) density[i] = 0;
    }
    if (i < sizeof(FrameCounters) / 4) reinterpret_cast<uint32_t *>(ctr)[i] = 0;
}

// VS_Kernel / GetNDCVertex (CCR/CUDARenderer.cu:673-700, CCR/CUDAROPs.cu:99-118).
__global__ __launch_bounds__(256) void k_vertex(const DevMesh *__restrict__ meshes,
                                                const uint32_t *__restrict__ block_mesh, FrameParams fp,
                                                float4 *__restrict__ pos4, float4 *__restrict__ attr,
                                                uint32_t *__restrict__ density)
{
    const uint32_t mi = block_mesh[blockIdx.x];
    const DevMesh &M = meshes[mi];
    const uint32_t v = (blockIdx.x - M.vs_block0) * 256u + threadIdx.x;
    if (v >= M.num_verts) return;
    const float4 *iv = reinterpret_cast<const float4 *>(M.ivert + (size_t)v * 16);
    const float4 q0 = iv[0], q1 = iv[1], q2 = iv[2], q3 = iv[3];
    const float px = q0.x, py = q0.y, pz = q0.z;
    const float u = q0.w, vv = q1.x;
    const float nx = q1.y, ny = q1.z, nz = q1.w;
    const float tx = q2.x, ty = q2.y, tz = q2.z;
    const float cr = q2.w, cg = q3.x, cb = q3.y;
    const float *W = M.wvp, *Wd = M.world;
    float o0 = m4(W, 0, 0) * px + m4(W, 0, 1) * py + m4(W, 0, 2) * pz + m4(W, 0, 3);
    float o1 = m4(W, 1, 0) * px + m4(W, 1, 1) * py + m4(W, 1, 2) * pz + m4(W, 1, 3);
    float o2 = m4(W, 2, 0) * px + m4(W, 2, 1) * py + m4(W, 2, 2) * pz + m4(W, 2, 3);
    const float o3 = m4(W, 3, 0) * px + m4(W, 3, 1) * py + m4(W, 3, 2) * pz + m4(W, 3, 3);
    o0 /= o3;
    o1 /= o3;
    o2 /= o3;
    const uint32_t g = M.vert_base + v;
    pos4[g] = make_float4(o0, o1, o2, o3);
    const float n0 = m4(Wd, 0, 0) * nx + m4(Wd, 0, 1) * ny + m4(Wd, 0, 2) * nz;
    const float n1 = m4(Wd, 1, 0) * nx + m4(Wd, 1, 1) * ny + m4(Wd, 1, 2) * nz;
    const float n2 = m4(Wd, 2, 0) * nx + m4(Wd, 2, 1) * ny + m4(Wd, 2, 2) * nz;
    const float t0 = m4(Wd, 0, 0) * tx + m4(Wd, 0, 1) * ty + m4(Wd, 0, 2) * tz;
    const float t1 = m4(Wd, 1, 0) * tx + m4(Wd, 1, 1) * ty + m4(Wd, 1, 2) * tz;
    const float t2 = m4(Wd, 2, 0) * tx + m4(Wd, 2, 1) * ty + m4(Wd, 2, 2) * tz;
    const float w0 = m4(Wd, 0, 0) * px + m4(Wd, 0, 1) * py + m4(Wd, 0, 2) * pz + m4(Wd, 0, 3);
    const float w1 = m4(Wd, 1, 0) * px + m4(Wd, 1, 1) * py + m4(Wd, 1, 2) * pz + m4(Wd, 1, 3);
    const float w2 = m4(Wd, 2, 0) * px + m4(Wd, 2, 1) * py + m4(Wd, 2, 2) * pz + m4(Wd, 2, 3);
    float vx = w0 - fp.cam_pos[0], vy = w1 - fp.cam_pos[1], vz = w2 - fp.cam_pos[2];
    normalize3(vx, vy, vz);
    float4 *a = attr + (size_t)g * 4;
    a[0] = make_float4(u, vv, n0, n1);
    a[1] = make_float4(n2, t0, t1, t2);
    a[2] = make_float4(vx, vy, vz, cr);
    a[3] = make_float4(cg, cb, 1.f / o3, 0.f);
    if (density) {
        // projected-vertex density feature of the capacity model (no parity role)
        const float sx = ((o0 + 1) / 2) * (float)fp.width;
        const float sy = ((1 - o1) / 2) * (float)fp.height;
        if (o3 > 0.f && sx >= 0.f && sy >= 0.f && sx < (float)fp.width && sy < (float)fp.height) {
            const uint32_t tile = ((uint32_t)sy / fp.tile) * fp.tiles_x + (uint32_t)sx / fp.tile;
            agg_increment(density, tile);
        }
    }
}

// TA_Kernel + TB_Kernel (CCR/CUDARenderer.cu:702-796,903-988) on the CPU
// path's rules (CCR/ERenderer.cpp:73-157, 603-624) with D4/D5.
template <bool SINGLE_PASS>
__global__ __launch_bounds__(256) void k_setup(const DevMesh *__restrict__ meshes,
                                               const uint32_t *__restrict__ block_mesh, FrameParams fp,
                                               const float4 *__restrict__ pos4, uint32_t *__restrict__ trirec,
                                               uint8_t *__restrict__ alive, uint32_t *__restrict__ tile_count,
                                               const uint32_t *__restrict__ tile_off,
                                               const uint32_t *__restrict__ tile_cap, uint32_t *__restrict__ queue,
                                               uint32_t queue_cap, uint32_t *__restrict__ survivors,
                                               uint32_t *__restrict__ big, uint32_t big_cap,
                                               uint2 *__restrict__ ovf, uint32_t ovf_cap,
                                               FrameCounters *__restrict__ ctr)
{
    const uint32_t mi = block_mesh[blockIdx.x];
    const DevMesh &M = meshes[mi];
    const uint32_t k = (blockIdx.x - M.tri_block0) * 256u + threadIdx.x;
    const bool valid = k < M.num_tris;
    const uint32_t gid = M.tri_base + k;
    bool ok = false;
    int xmin = 0, ymin = 0, xmax = 0, ymax = 0;
    if (valid) {
        uint32_t i0, i1, i2;
        if (M.topology == 0) {
            i0 = M.idx[3 * (size_t)k]; i1 = M.idx[3 * (size_t)k + 1]; i2 = M.idx[3 * (size_t)k + 2];
        } else {  // strip: odd triangles swap their last two indices
            const bool odd = (k & 1u) != 0;
            i0 = M.idx[k];
            i1 = odd ? M.idx[k + 2] : M.idx[k + 1];
            i2 = odd ? M.idx[k + 1] : M.idx[k + 2];
        }
        i0 += M.vert_base; i1 += M.vert_base; i2 += M.vert_base;
        const float4 P0 = pos4[i0], P1 = pos4[i1], P2 = pos4[i2];
        ok = true;
        if (fp.cull_mode != 2) {  // back / front face (CCR/ERenderer.cpp:82-101)
            const float e1x = P1.x - P0.x, e1y = P1.y - P0.y, e1z = P1.z - P0.z;
            const float e2x = P2.x - P0.x, e2y = P2.y - P0.y, e2z = P2.z - P0.z;
            float fx = e1y * e2z - e1z * e2y;
            float fy = e1z * e2x - e1x * e2z;
            float fz = e1x * e2y - e1y * e2x;
            normalize3(fx, fy, fz);
            const float cv = fp.cam_fwd[0] * fx + fp.cam_fwd[1] * fy + fp.cam_fwd[2] * fz;
            ok = fp.cull_mode == 0 ? !(cv <= 0.f) : !(cv >= 0.f);
        }
        // IsTriangleVisible (CCR/CUDAROPs.cu:147-178)
        if (ok && (all_outside(P0.x, P1.x, P2.x, -1.f, 1.f) || all_outside(P0.y, P1.y, P2.y, -1.f, 1.f) ||
                   all_outside(P0.z, P1.z, P2.z, 0.f, 1.f)))
            ok = false;
        if (ok) {
            const float W = (float)fp.width, H = (float)fp.height;
            // NDCToScreenSpace (CCR/CUDAROPs.cu:88-97)
            const float sx0 = ((P0.x + 1) / 2) * W, sy0 = ((1 - P0.y) / 2) * H;
            const float sx1 = ((P1.x + 1) / 2) * W, sy1 = ((1 - P1.y) / 2) * H;
            const float sx2 = ((P2.x + 1) / 2) * W, sy2 = ((1 - P2.y) / 2) * H;
            // Cross(v0 - v1, v0 - v2) (CCR/CUDAROPs.cu:336)
            const float ex = sx0 - sx1, ey = sy0 - sy1, fx2 = sx0 - sx2, fy2 = sy0 - sy2;
            const float cr = ex * fy2 - ey * fx2;
            if (!(cr != 0.f) || !isfinite(cr)) ok = false;
            if (ok) {
                const float area = fabsf(cr);
                // GetBoundingBox with D4 (CCR/CUDAROPs.cu:180-200)
                float mnx = sx0, mxx = sx0, mny = sy0, mxy = sy0;
                if (sx1 < mnx) mnx = sx1;
                if (sx2 < mnx) mnx = sx2;
                if (sx1 > mxx) mxx = sx1;
                if (sx2 > mxx) mxx = sx2;
                if (sy1 < mny) mny = sy1;
                if (sy2 < mny) mny = sy2;
                if (sy1 > mxy) mxy = sy1;
                if (sy2 > mxy) mxy = sy2;
                xmin = clamp_trunc(mnx, W + 2.f) - 1;
                ymin = clamp_trunc(mny, H + 2.f) - 1;
                xmax = clamp_trunc(mxx, W + 2.f) + 1;
                ymax = clamp_trunc(mxy, H + 2.f) + 1;
                xmin = xmin < 0 ? 0 : xmin;
                ymin = ymin < 0 ? 0 : ymin;
                xmax = xmax > (int)fp.width ? (int)fp.width : xmax;
                ymax = ymax > (int)fp.height ? (int)fp.height : ymax;
                uint4 *r = reinterpret_cast<uint4 *>(trirec + (size_t)gid * kTriRecWords);
                r[0] = make_uint4(__float_as_uint(sx0), __float_as_uint(sy0), __float_as_uint(sx1),
                                  __float_as_uint(sy1));
                r[1] = make_uint4(__float_as_uint(sx2), __float_as_uint(sy2), __float_as_uint(P0.z),
                                  __float_as_uint(P1.z));
                r[2] = make_uint4(__float_as_uint(P2.z), __float_as_uint(area),
                                  (uint32_t)xmin | ((uint32_t)ymin << 16), (uint32_t)xmax | ((uint32_t)ymax << 16));
                r[3] = make_uint4(i0, i1, i2, mi);
            }
        }
        alive[gid] = ok ? 1 : 0;
    }
    const uint64_t am = __ballot(ok);
    if (lane_id() == 0 && am) atomicAdd(&ctr->num_alive, (uint32_t)__popcll(am));

    // binning: tile rectangle of the pixel range [xmin,xmax) x [ymin,ymax)
    const bool nonempty = ok && xmin < xmax && ymin < ymax;
    uint32_t tx0 = 0, tx1 = 0, ty0 = 0, ty1 = 0, span = 0;
    if (nonempty) {
        tx0 = (uint32_t)xmin / fp.tile; tx1 = (uint32_t)(xmax - 1) / fp.tile;
        ty0 = (uint32_t)ymin / fp.tile; ty1 = (uint32_t)(ymax - 1) / fp.tile;
        span = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
    }
    const bool is_big = nonempty && span > fp.big_threshold;
    const bool binned = nonempty && !is_big;
    const uint32_t bslot = wave_append(&ctr->num_big, is_big);
    if (is_big && bslot < big_cap) big[bslot] = gid;
    if (is_big && bslot >= big_cap) ctr->storage_overflow = 1;
    const uint32_t sslot = wave_append(&ctr->num_binned, binned);
    if (binned) survivors[sslot] = gid;
    if (binned) {
        const uint32_t w = tx1 - tx0 + 1;
        for (uint32_t s = 0; s < span; ++s) {
            const uint32_t t = (ty0 + s / w) * fp.tiles_x + tx0 + s % w;
            const uint32_t slot = agg_increment(tile_count, t);
            if (SINGLE_PASS) {
                if (slot < tile_cap[t]) {
                    const uint32_t pos = tile_off[t] + slot;
                    if (pos < queue_cap) queue[pos] = gid;
                } else {
                    // capacity exceeded: keep the entry for the exact fallback
                    const uint32_t o = atomicAdd(&ctr->ovf_entries, 1u);
                    if (o < ovf_cap) ovf[o] = make_uint2(t, gid);
                    else ctr->storage_overflow = 1;
                }
            }
        }
    }
}

// Single-workgroup exclusive scan (n up to a few hundred thousand tiles).
// out_off[n] = total.  is_capacity: in = capacities, the total is also the
// capacity sum; otherwise in = exact counts and the total is the pair count.
__global__ __launch_bounds__(1024) void k_scan(uint32_t n, const uint32_t *__restrict__ in,
                                               uint32_t *__restrict__ out_off, uint32_t *__restrict__ zero_cursor,
                                               uint32_t limit, FrameCounters *__restrict__ ctr, int is_capacity)
{
    __shared__ uint32_t part[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t b = tid * per;
    const uint32_t e = b + per < n ? b + per : n;
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += in[i];
    part[tid] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t run = tid ? part[tid - 1] : 0;
    for (uint32_t i = b; i < e; ++i) {
        out_off[i] = run;
        run += in[i];
        if (zero_cursor) zero_cursor[i] = 0;
    }
    if (tid == 1023) {
        out_off[n] = part[1023];
        if (is_capacity) ctr->capacity_total = part[1023];
        else ctr->total_pairs = part[1023];
        if (part[1023] > limit) ctr->storage_overflow = 1;
    }
}

// Exact second pass over the binned survivors (the order inside a queue is
// irrelevant: the raster resolves depth order-independently).
__global__ __launch_bounds__(256) void k_fill(FrameParams fp, const uint32_t *__restrict__ survivors,
                                              const uint32_t *__restrict__ trirec,
                                              uint32_t *__restrict__ cursor, const uint32_t *__restrict__ tile_off,
                                              uint32_t *__restrict__ queue, uint32_t queue_cap,
                                              const FrameCounters *__restrict__ ctr)
{
    const uint32_t n = ctr->num_binned;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i - threadIdx.x < n; i += gridDim.x * 256u) {
        if (i < n) {
            const uint32_t gid = survivors[i];
            const uint2 bb = *reinterpret_cast<const uint2 *>(trirec + (size_t)gid * kTriRecWords + 10);
            const uint32_t xmin = bb.x & 0xFFFFu, ymin = bb.x >> 16, xmax = bb.y & 0xFFFFu, ymax = bb.y >> 16;
            const uint32_t tx0 = xmin / fp.tile, tx1 = (xmax - 1) / fp.tile;
            const uint32_t ty0 = ymin / fp.tile, ty1 = (ymax - 1) / fp.tile;
            const uint32_t w = tx1 - tx0 + 1, span = w * (ty1 - ty0 + 1);
            for (uint32_t s = 0; s < span; ++s) {
                const uint32_t t = (ty0 + s / w) * fp.tiles_x + tx0 + s % w;
                const uint32_t slot = agg_increment(cursor, t);
                const uint32_t pos = tile_off[t] + slot;
                if (pos < queue_cap) queue[pos] = gid;
            }
        }
    }
}

// Single-pass fallback, step 1 (per tile): the exact count of a tile is its
// final fill cursor; the entries beyond its capacity get a pool region.
__global__ __launch_bounds__(256) void k_fallback_alloc(uint32_t n, const uint32_t *__restrict__ count,
                                                        const uint32_t *__restrict__ cap,
                                                        uint32_t *__restrict__ tile_pool,
                                                        uint32_t *__restrict__ fb_cursor, uint32_t pool_cap,
                                                        FrameCounters *__restrict__ ctr)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    uint32_t c = 0, excess = 0;
    if (t < n) {
        c = count[t];
        excess = c > cap[t] ? c - cap[t] : 0;
    }
    uint32_t off = 0;
    if (excess) {
        off = atomicAdd(&ctr->pool_used, excess);
        if (off + excess > pool_cap) ctr->storage_overflow = 1;
    }
    if (t < n) {
        tile_pool[t] = excess ? off : 0xFFFFFFFFu;
        fb_cursor[t] = 0;
    }
    // frame totals: exact pair count and number of overflowed tiles
    uint32_t tot = c, ovt = excess ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) {
        tot += __shfl_xor(tot, o);
        ovt += __shfl_xor(ovt, o);
    }
    if (lane_id() == 0) {
        if (tot) atomicAdd(&ctr->total_pairs, tot);
        if (ovt) atomicAdd(&ctr->overflow_tiles, ovt);
    }
}

// Step 2 (per overflow entry): exact placement of the entries that did not fit.
__global__ __launch_bounds__(256) void k_fallback_fill(const uint2 *__restrict__ ovf, uint32_t ovf_cap,
                                                       const uint32_t *__restrict__ tile_pool,
                                                       uint32_t *__restrict__ fb_cursor, uint32_t *__restrict__ pool,
                                                       uint32_t pool_cap, const FrameCounters *__restrict__ ctr)
{
    uint32_t n = ctr->ovf_entries;
    n = n < ovf_cap ? n : ovf_cap;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const uint2 e = ovf[i];
        const uint32_t po = tile_pool[e.x];
        if (po == 0xFFFFFFFFu) continue;
        const uint32_t pos = po + atomicAdd(fb_cursor + e.x, 1u);
        if (pos < pool_cap) pool[pos] = e.y;
    }
}

// RA_PerTileKernel (CCR/CUDARenderer.cu:1039-1129) redesigned: one workgroup
// per tile, depth + visibility resolved in LDS with 64-bit (z, ~id) keys so
// the queue order does not matter.  Small footprints are rasterized one
// triangle per lane (LDS atomic min); larger ones are staged in LDS and
// rasterized one pixel per lane.
template <int T, int NT>
__global__ __launch_bounds__(NT) void k_raster(FrameParams fp, const uint32_t *__restrict__ tile_count,
                                               const uint32_t *__restrict__ tile_off,
                                               const uint32_t *__restrict__ tile_cap,
                                               const uint32_t *__restrict__ tile_pool,
                                               const uint32_t *__restrict__ queue, uint32_t queue_cap,
                                               const uint32_t *__restrict__ pool, uint32_t pool_cap,
                                               const uint32_t *__restrict__ big, uint32_t big_cap,
                                               const FrameCounters *__restrict__ ctr,
                                               const uint32_t *__restrict__ trirec, uint32_t *__restrict__ vis,
                                               float *__restrict__ depth, float *__restrict__ tile_min_depth,
                                               uint32_t small_limit, uint32_t grid_tiles)
{
    constexpr int NPIX = T * T;
    constexpr int PPT = NPIX / NT;
    static_assert(NPIX % NT == 0, "tile pixels must be a multiple of the workgroup size");
    __shared__ uint64_t key[NPIX];
    __shared__ TriLds stage[NT];
    __shared__ uint32_t nstage[2];

    const uint32_t ntiles = fp.tiles_x * fp.tiles_y;
    // XCD-aware order: consecutive workgroups go round-robin over the 8 XCDs,
    // so give each XCD a contiguous band of tiles (neighbours share triangles).
    const uint32_t per_xcd = grid_tiles / 8u;
    const uint32_t tile = (blockIdx.x % 8u) * per_xcd + blockIdx.x / 8u;
    if (tile >= ntiles) return;
    const uint32_t tx = tile % fp.tiles_x, ty = tile / fp.tiles_x;
    const int X0 = (int)(tx * T), Y0 = (int)(ty * T);
    const uint32_t tid = threadIdx.x;

    for (int p = tid; p < NPIX; p += NT) key[p] = kClearKey;
    if (tid == 0) { nstage[0] = 0; nstage[1] = 0; }

    // list A: the tile's queue region; list B: its fallback pool region
    const uint32_t cnt = tile_count[tile];
    uint32_t n_a = cnt, n_b = 0;
    const uint32_t off = tile_off[tile];
    const uint32_t *list_a = queue + off;
    const uint32_t *list_b = pool;
    if (tile_cap) {
        const uint32_t cp = tile_cap[tile];
        if (cnt > cp) {
            n_a = cp;
            const uint32_t po = tile_pool[tile];
            if (po != 0xFFFFFFFFu) {
                list_b = pool + po;
                const uint32_t room_b = po < pool_cap ? pool_cap - po : 0;
                n_b = cnt - cp;
                n_b = n_b < room_b ? n_b : room_b;
            }
        }
    }
    const uint32_t room_a = off < queue_cap ? queue_cap - off : 0;
    n_a = n_a < room_a ? n_a : room_a;
    const uint32_t n_tile = n_a + n_b;
    uint32_t n_big = ctr->num_big;
    n_big = n_big < big_cap ? n_big : big_cap;
    const uint32_t total = n_tile + n_big;
    const uint32_t num_tris = fp.num_tris;
    __syncthreads();

    // rows of the tile owned by this wave in the pixel-parallel phase
    const int wave = (int)(tid / 64u);
    uint32_t parity = 0;
    for (uint32_t base = 0; base < total; base += NT, parity ^= 1u) {
        const uint32_t i = base + tid;
        if (i < total) {
            const uint32_t id = i < n_a ? list_a[i] : (i < n_tile ? list_b[i - n_a] : big[i - n_tile]);
            if (id < num_tris) {
                TriLds t;
                load_tri(trirec, id, t);
                const int bx0 = (int)(t.xr & 0xFFFFu), by0 = (int)(t.xr >> 16);
                const int bx1 = (int)(t.yr & 0xFFFFu), by1 = (int)(t.yr >> 16);
                const int x0 = bx0 > X0 ? bx0 : X0, x1 = bx1 < X0 + T ? bx1 : X0 + T;
                const int y0 = by0 > Y0 ? by0 : Y0, y1 = by1 < Y0 + T ? by1 : Y0 + T;
                if (x0 < x1 && y0 < y1) {
                    const uint32_t npix = (uint32_t)((x1 - x0) * (y1 - y0));
                    if (npix <= small_limit) {
                        for (int py = y0; py < y1; ++py)
                            for (int px = x0; px < x1; ++px) {
                                float z;
                                if (fragment(t, (float)px, (float)py, z))
                                    atomicMin((unsigned long long *)&key[(py - Y0) * T + (px - X0)],
                                              (unsigned long long)make_key(z, id));
                            }
                    } else {
                        t.xr = (uint32_t)x0 | ((uint32_t)x1 << 16);
                        t.yr = (uint32_t)y0 | ((uint32_t)y1 << 16);
                        const uint32_t slot = atomicAdd(&nstage[parity], 1u);
                        stage[slot] = t;
                    }
                }
            }
        }
        __syncthreads();
        const uint32_t ns = nstage[parity];
        if (tid == 0) nstage[parity ^ 1u] = 0;
        if (ns) {
            uint64_t best[PPT];
            int pxs[PPT], pys[PPT];
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int p = (int)tid + q * NT;
                pxs[q] = X0 + p % T;
                pys[q] = Y0 + p / T;
                best[q] = key[p];
            }
            // wave-uniform row range of this wave's pixels
            const int wy0 = Y0 + (wave * 64) / T;
            const int wy1 = Y0 + ((wave * 64) + (PPT - 1) * NT + 63) / T + 1;
            for (uint32_t j = 0; j < ns; ++j) {
                const TriLds &t = stage[j];
                const int x0 = (int)(t.xr & 0xFFFFu), x1 = (int)(t.xr >> 16);
                const int y0 = (int)(t.yr & 0xFFFFu), y1 = (int)(t.yr >> 16);
                if (PPT == 1 && (y1 <= wy0 || y0 >= wy1)) continue;
#pragma unroll
                for (int q = 0; q < PPT; ++q) {
                    const int px = pxs[q], py = pys[q];
                    if (px >= x0 && px < x1 && py >= y0 && py < y1) {
                        float z;
                        if (fragment(t, (float)px, (float)py, z)) {
                            const uint64_t kk = make_key(z, t.id);
                            best[q] = kk < best[q] ? kk : best[q];
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < PPT; ++q) key[(int)tid + q * NT] = best[q];
        }
        __syncthreads();
    }

    // write out (coalesced rows of T pixels)
    float dmin = 1.0f;
    for (int p = tid; p < NPIX; p += NT) {
        const int px = X0 + p % T, py = Y0 + p / T;
        const uint64_t kk = key[p];
        const float d = __uint_as_float((uint32_t)(kk >> 32));
        dmin = d < dmin ? d : dmin;
        if (px < (int)fp.width && py < (int)fp.height) {
            const size_t o = (size_t)py * fp.width + px;
            vis[o] = ~(uint32_t)kk - 1u;
            depth[o] = d;
        }
    }
    if (tile_min_depth) {
        // per-tile min depth (feature of the capacity model); wave min then LDS
        for (int off = 32; off > 0; off >>= 1) {
            const float o = __shfl_xor(dmin, off);
            dmin = o < dmin ? o : dmin;
        }
        __shared__ float wmin[NT / 64 > 0 ? NT / 64 : 1];
        if ((tid & 63u) == 0) wmin[tid / 64u] = dmin;
        __syncthreads();
        if (tid == 0) {
            float m = wmin[0];
            for (int w = 1; w < NT / 64; ++w) m = wmin[w] < m ? wmin[w] : m;
            tile_min_depth[tile] = m;
        }
    }
}

// ------------------------------------------------------------- shading

__device__ __forceinline__ uint32_t pack_channel(float c)
{
    const float v = c * 255.f;
    if (!(v > 0.f)) return 0u;
    if (v >= 255.f) return 255u;
    return (uint32_t)v;
}

// GetRGBAFromColour (CCR/RGBRaw.h:23-32) in memory order R,G,B,A=255 with
// the saturating float->u8 conversion of the reference GPU.
__device__ __forceinline__ uint32_t pack_rgba(float r, float g, float b)
{
    return pack_channel(r) | (pack_channel(g) << 8) | (pack_channel(b) << 16) | 0xFF000000u;
}

__device__ __forceinline__ bool texel_coord(float uv, uint32_t n, int &out)
{
    const float f = uv * (float)n + 0.5f;
    if (!(f > -1.f && f < (float)n)) return false;
    out = (int)f;
    return true;
}

// CUDATextureSampler::SamplePoint / SampleLinear (CCR/CUDATextureSampler.cuh:86-190).
__device__ __forceinline__ void sample(const DevTexture &t, float u, float v, int mode, float &r, float &g,
                                       float &b)
{
    int x, y;
    if (!texel_coord(u, t.w, x) || !texel_coord(v, t.h, y)) { r = 1.f; g = 0.f; b = 1.f; return; }
    const int w = (int)t.w;
    const int orig = x + y * w;
    if (mode == 0) {
        const uint32_t s = t.texels[orig];
        r = (float)(s & 0xFFu) / 255.f;
        g = (float)((s >> 8) & 0xFFu) / 255.f;
        b = (float)((s >> 16) & 0xFFu) / 255.f;
        return;
    }
    const int texSize = w * (int)t.h;
    int nb[4];
    nb[0] = orig - 1; if (nb[0] < 0) nb[0] = orig;
    nb[1] = orig + 1; if (nb[1] >= texSize) nb[1] = orig;
    nb[2] = orig + w; if (nb[2] >= texSize) nb[2] = orig;
    nb[3] = orig - w; if (nb[3] < 0) nb[3] = orig;
    const float weight = 1.f / (4 * 2);
    float ar = 0.f, ag = 0.f, ab = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t s = t.texels[nb[i]];
        ar += (float)(s & 0xFFu) * weight;
        ag += (float)((s >> 8) & 0xFFu) * weight;
        ab += (float)((s >> 16) & 0xFFu) * weight;
    }
    const uint32_t s = t.texels[orig];
    ar += (float)(s & 0xFFu) / 2.f;
    ag += (float)((s >> 8) & 0xFFu) / 2.f;
    ab += (float)((s >> 16) & 0xFFu) / 2.f;
    const float rev = 1.0f / 255.f;
    r = ar * rev; g = ag * rev; b = ab * rev;
}

// PS_Kernel + ShadePixelUnsafe (CCR/CUDARenderer.cu:833-897, CCR/CUDAROPs.cu:472-514)
// with the attribute interpolation of RasterizePixelAtomic (CCR/CUDAROPs.cu:347-380).
__global__ __launch_bounds__(256) void k_shade(FrameParams fp, const DevMesh *__restrict__ meshes,
                                               const uint32_t *__restrict__ vis, const float *__restrict__ depth,
                                               const uint32_t *__restrict__ trirec,
                                               const float4 *__restrict__ attr, uint32_t *__restrict__ rgba)
{
    const uint32_t x = blockIdx.x * 256u + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (x >= fp.width) return;
    const size_t pid = (size_t)y * fp.width + x;
    const uint32_t id = vis[pid];
    if (id >= fp.num_tris) {
        rgba[pid] = fp.visualisation == 0 ? fp.clear_rgba : 0xFF000000u;
        return;
    }
    const uint4 *rp = reinterpret_cast<const uint4 *>(trirec + (size_t)id * kTriRecWords);
    const uint4 q0 = rp[0], q1 = rp[1], q2 = rp[2], q3 = rp[3];
    const float sx0 = __uint_as_float(q0.x), sy0 = __uint_as_float(q0.y);
    const float sx1 = __uint_as_float(q0.z), sy1 = __uint_as_float(q0.w);
    const float sx2 = __uint_as_float(q1.x), sy2 = __uint_as_float(q1.y);
    const float area = __uint_as_float(q2.y);
    const float px = (float)x, py = (float)y;
    float w2 = (px - sx0) * (sy1 - sy0) - (py - sy0) * (sx1 - sx0);
    float w0 = (px - sx1) * (sy2 - sy1) - (py - sy1) * (sx2 - sx1);
    float w1 = (px - sx2) * (sy0 - sy2) - (py - sy2) * (sx0 - sx2);
    w0 = w0 / area; w1 = w1 / area; w2 = w2 / area;
    const float4 *A0 = attr + (size_t)q3.x * 4, *A1 = attr + (size_t)q3.y * 4, *A2 = attr + (size_t)q3.z * 4;
    const float4 a00 = A0[0], a01 = A0[1], a02 = A0[2], a03 = A0[3];
    const float4 a10 = A1[0], a11 = A1[1], a12 = A1[2], a13 = A1[3];
    const float4 a20 = A2[0], a21 = A2[1], a22 = A2[2], a23 = A2[3];
    const float i0 = a03.z, i1 = a13.z, i2 = a23.z;
    const float wI = 1.f / (i0 * w0 + i1 * w1 + i2 * w2);
#define RAST_INTERP(c0, c1, c2) (w0 * ((c0) * i0) + w1 * ((c1) * i1) + w2 * ((c2) * i2))
    const float u = RAST_INTERP(a00.x, a10.x, a20.x) * wI;
    const float v = RAST_INTERP(a00.y, a10.y, a20.y) * wI;
    const float nx = RAST_INTERP(a00.z, a10.z, a20.z) * wI;
    const float ny = RAST_INTERP(a00.w, a10.w, a20.w) * wI;
    const float nz = RAST_INTERP(a01.x, a11.x, a21.x) * wI;
    if (fp.visualisation == 1) {  // depth colour, Remap(z, 0.985, 1)
        rgba[pid] = pack_rgba((depth[pid] - 0.985f) / (1.f - 0.985f), 0.f, 0.f);
        return;
    }
    if (fp.visualisation == 2) {  // normal colour
        rgba[pid] = pack_rgba(nx, ny, nz);
        return;
    }
    const DevMesh &M = meshes[q3.w < fp.num_meshes ? q3.w : 0];
    if (!M.has_diffuse) {  // vertex colour (perspective-weighted, not x w)
        rgba[pid] = pack_rgba(RAST_INTERP(a02.w, a12.w, a22.w), RAST_INTERP(a03.x, a13.x, a23.x),
                              RAST_INTERP(a03.y, a13.y, a23.y));
        return;
    }
    const float tx = RAST_INTERP(a01.y, a11.y, a21.y);
    const float ty = RAST_INTERP(a01.z, a11.z, a21.z);
    const float tz = RAST_INTERP(a01.w, a11.w, a21.w);
    float vx = RAST_INTERP(a02.x, a12.x, a22.x);
    float vy = RAST_INTERP(a02.y, a12.y, a22.y);
    float vz = RAST_INTERP(a02.z, a12.z, a22.z);
#undef RAST_INTERP
    normalize3(vx, vy, vz);
    float dr, dg, db, nr, ng, nb, sr, sg, sb, gr, gg, gb;
    sample(M.tex[0], u, v, fp.sample_state, dr, dg, db);
    sample(M.tex[1], u, v, fp.sample_state, nr, ng, nb);
    sample(M.tex[2], u, v, fp.sample_state, sr, sg, sb);
    sample(M.tex[3], u, v, fp.sample_state, gr, gg, gb);
    (void)gg; (void)gb;
    // binormal = -Cross(tan, n); N = {tan, binormal, n} * (2 * sample - 1)
    const float cx = ty * nz - tz * ny, cy = tz * nx - tx * nz, cz = tx * ny - ty * nx;
    const float bx = -cx, by = -cy, bz = -cz;
    const float fx = 2.f * nr - 1.f, fy = 2.f * ng - 1.f, fz = 2.f * nb - 1.f;
    const float Nx = tx * fx + bx * fy + nx * fz;
    const float Ny = ty * fx + by * fy + ny * fz;
    const float Nz = tz * fx + bz * fy + nz * fz;
    const float Lx = 0.577f, Ly = -0.577f, Lz = -0.577f;
    float oa = (-Nx) * Lx + (-Ny) * Ly + (-Nz) * Lz;  // ClampFast(x, 0, x) leaves x unchanged
    oa /= 3.14159265358979323846f;
    oa *= 7.0f;
    const float twoDot = 2.f * (Lx * Nx + Ly * Ny + Lz * Nz);
    const float rx = -Lx + twoDot * Nx, ry = -Ly + twoDot * Ny, rz = -Lz + twoDot * Nz;
    float angle = vx * rx + vy * ry + vz * rz;
    angle = angle < 0.f ? 0.f : angle;
    angle = angle > 1.f ? 1.f : angle;
    angle = powf(angle, gr * 25.f);
    float outr = 0.05f + dr * oa + sr * angle;
    float outg = 0.05f + dg * oa + sg * angle;
    float outb = 0.05f + db * oa + sb * angle;
    const float mgb = outg < outb ? outb : outg;
    const float mx = outr < mgb ? mgb : outr;
    if (mx > 1.f) {
        const float rev = 1.0f / mx;
        outr *= rev; outg *= rev; outb *= rev;
    }
    rgba[pid] = pack_rgba(outr, outg, outb);
}

inline uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

hipError_t launch_frame_init(const FrameParams &fp, const FrameBuffers &b, bool single_pass, hipStream_t s)
{
    const uint32_t n = fp.tiles_x * fp.tiles_y;
    const uint32_t nn = n > 64 ? n : 64;
    hipLaunchKernelGGL(k_frame_init, dim3(ceil_div(nn, 256)), dim3(256), 0, s, b.tile_count,
                       single_pass ? b.tile_pool : nullptr, n, b.counters, b.density);
    return hipGetLastError();
}

hipError_t launch_vertex(const FrameParams &fp, const FrameBuffers &b, hipStream_t s)
{
    if (b.vs_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_vertex, dim3(b.vs_blocks), dim3(256), 0, s, b.meshes, b.vs_block_mesh, fp, b.pos4, b.attr,
                       b.density);
    return hipGetLastError();
}

hipError_t launch_setup(const FrameParams &fp, const FrameBuffers &b, bool single_pass, hipStream_t s)
{
    if (b.tri_blocks == 0) return hipSuccess;
    if (single_pass)
        hipLaunchKernelGGL(k_setup<true>, dim3(b.tri_blocks), dim3(256), 0, s, b.meshes, b.tri_block_mesh, fp, b.pos4,
                           b.trirec, b.alive, b.tile_count, b.tile_off, b.tile_cap, b.queue, b.queue_cap,
                           b.survivors, b.big, b.big_cap, b.ovf, b.ovf_cap, b.counters);
    else
        hipLaunchKernelGGL(k_setup<false>, dim3(b.tri_blocks), dim3(256), 0, s, b.meshes, b.tri_block_mesh, fp,
                           b.pos4, b.trirec, b.alive, b.tile_count, b.tile_off, b.tile_cap, b.queue, b.queue_cap,
                           b.survivors, b.big, b.big_cap, b.ovf, b.ovf_cap, b.counters);
    return hipGetLastError();
}

hipError_t launch_scan(uint32_t n, const uint32_t *in, uint32_t *out_off, uint32_t *zero_cursor, uint32_t limit,
                       FrameCounters *counters, bool is_capacity, hipStream_t s)
{
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, n, in, out_off, zero_cursor, limit, counters,
                       is_capacity ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_fill(const FrameParams &fp, const FrameBuffers &b, hipStream_t s)
{
    const uint32_t blocks = ceil_div(fp.num_tris > 0 ? fp.num_tris : 1, 256);
    hipLaunchKernelGGL(k_fill, dim3(blocks < 2048 ? blocks : 2048), dim3(256), 0, s, fp, b.survivors, b.trirec,
                       b.tile_cursor, b.tile_off, b.queue, b.queue_cap, b.counters);
    return hipGetLastError();
}

hipError_t launch_fallback(const FrameParams &fp, const FrameBuffers &b, hipStream_t s)
{
    const uint32_t n = fp.tiles_x * fp.tiles_y;
    hipLaunchKernelGGL(k_fallback_alloc, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, b.tile_count, b.tile_cap,
                       b.tile_pool, b.tile_cursor, b.pool_cap, b.counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fallback_fill, dim3(256), dim3(256), 0, s, b.ovf, b.ovf_cap, b.tile_pool, b.tile_cursor,
                       b.pool, b.pool_cap, b.counters);
    return hipGetLastError();
}

hipError_t launch_raster(const FrameParams &fp, const FrameBuffers &b, uint32_t small_limit, hipStream_t s)
{
    const uint32_t n = fp.tiles_x * fp.tiles_y;
    const uint32_t grid = ceil_div(n, 8) * 8;
    switch (fp.tile) {
    case 8:
        hipLaunchKernelGGL((k_raster<8, 64>), dim3(grid), dim3(64), 0, s, fp, b.tile_count, b.tile_off, b.tile_cap,
                           b.tile_pool, b.queue, b.queue_cap, b.pool, b.pool_cap, b.big, b.big_cap, b.counters, b.trirec, b.vis,
                           b.depth, b.tile_min_depth, small_limit, grid);
        break;
    case 16:
        hipLaunchKernelGGL((k_raster<16, 256>), dim3(grid), dim3(256), 0, s, fp, b.tile_count, b.tile_off,
                           b.tile_cap, b.tile_pool, b.queue, b.queue_cap, b.pool, b.pool_cap, b.big, b.big_cap, b.counters,
                           b.trirec, b.vis, b.depth, b.tile_min_depth, small_limit, grid);
        break;
    case 32:
        hipLaunchKernelGGL((k_raster<32, 256>), dim3(grid), dim3(256), 0, s, fp, b.tile_count, b.tile_off,
                           b.tile_cap, b.tile_pool, b.queue, b.queue_cap, b.pool, b.pool_cap, b.big, b.big_cap, b.counters,
                           b.trirec, b.vis, b.depth, b.tile_min_depth, small_limit, grid);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_shade(const FrameParams &fp, const FrameBuffers &b, hipStream_t s)
{
    hipLaunchKernelGGL(k_shade, dim3(ceil_div(fp.width, 256), fp.height), dim3(256), 0, s, fp, b.meshes, b.vis,
                       b.depth, b.trirec, b.attr, b.rgba);
    return hipGetLastError();
}

}  // namespace rast
