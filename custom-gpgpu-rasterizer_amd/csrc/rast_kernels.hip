// rast_kernels.hip -- the per-frame rasterization kernels for gfx950.
//
// Built with -ffp-contract=off and no fast-math: every float expression below
// is evaluated as written (one IEEE binary32 operation each, left to right),
// division and sqrt correctly rounded, so visibility and depth are bitwise the
// CPU restatement of the reference path (the shading's lighting terms, held to
// RGBA8 within 1/255, use explicit fused multiply-adds and the hardware
// reciprocal, rsqrt, exp2 and log2).  Reference files are cited relative
// to kdedeurw/Custom-GPGPU-Rasterizer (CCR = CustomCUDARasterizer/).
//
// Depth resolution is order-free: every fragment becomes a 64-bit key
// (z bits << 32 | ~(id + 1)) and the smallest key wins, which is exactly the
// sequential "reject if z > depth, later triangle wins ties" rule of
// CCR/ERenderer.cpp:544-554 applied in ascending triangle order.
//
// Pipeline (one launch per stage for all jobs and meshes of a launch):
//   k_vertex   per vertex: clip/NDC position + shading attributes; beside
//              it, single-pass jobs' tile capacities and queue regions, and
//              the reset of every job's per-tile counts (the frame counters
//              start at zero with the launch's staging upload)
//   k_setup    per triangle: assembly, cull, screen setup record; triangles
//              whose bounding box has at most micro_limit pixels are
//              rasterized on the spot with 64-bit atomic min into the key
//              buffer; larger ones are counted per tile (exact two-pass) or
//              written straight into per-tile queues (single pass); triangles
//              spanning more than big_threshold tiles go to one screen list
//   k_scan     per-tile exclusive scan of the exact counts
//   k_fill     exact second pass over the binned triangles
//   k_resolve  one workgroup per tile: keys -> LDS (and cleared for the next
//              frame), queued triangles rasterized in LDS (single pass: a
//              tile whose count exceeded its capacity first gathers its
//              remaining entries from the overflow list), vis/depth written
//   k_shade    per pixel, deferred from the visibility buffer; one instance per
//              sampling mode; its first block of each job publishes the
//              frame counters to the host (k_publish for unshaded jobs)
#include <hip/hip_runtime.h>

#include "rast_capacity.h"
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

// Global-memory access through global-address-space pointers.  Pointers
// read from the job and mesh tables are generic to the compiler, which would
// emit flat loads, stores and atomics; those also count against the LDS /
// scalar wait counter, so every LDS access behind them would wait for the
// global operations in flight (the rasterizer loops mix both).
template <typename T>
__device__ __forceinline__ T ldg(const T *p)
{
    return *(const __attribute__((address_space(1))) T *)p;
}

// uint4 / float4 are classes whose copy takes a generic reference (a flat
// load again): load them as the compiler's own 4-vectors.
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint2 ldg(const uint2 *p)
{
    const u32x2_t v = *(const __attribute__((address_space(1))) u32x2_t *)p;
    return make_uint2(v.x, v.y);
}

__device__ __forceinline__ uint4 ldg(const uint4 *p)
{
    const u32x4_t v = *(const __attribute__((address_space(1))) u32x4_t *)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ float4 ldg(const float4 *p)
{
    const f32x4_t v = *(const __attribute__((address_space(1))) f32x4_t *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ DevTexture ldg(const DevTexture *p)
{
    static_assert(sizeof(DevTexture) == 16, "texture descriptor: pointer, width, height");
    const u32x4_t v = *(const __attribute__((address_space(1))) u32x4_t *)p;
    DevTexture t;
    t.texels = reinterpret_cast<const uint32_t *>(((uint64_t)v.y << 32) | v.x);
    t.w = v.z;
    t.h = v.w;
    return t;
}

template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gptr(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

template <typename T>
__device__ __forceinline__ void stg(T *p, T v)
{
    *gptr(p) = v;
}

__device__ __forceinline__ void stg(uint4 *p, uint4 v)
{
    *(__attribute__((address_space(1))) u32x4_t *)p = u32x4_t{ v.x, v.y, v.z, v.w };
}

__device__ __forceinline__ void stg(float4 *p, float4 v)
{
    *(__attribute__((address_space(1))) f32x4_t *)p = f32x4_t{ v.x, v.y, v.z, v.w };
}

__device__ __forceinline__ void stg(uint2 *p, uint2 v)
{
    typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
    *(__attribute__((address_space(1))) u32x2_t *)p = u32x2_t{ v.x, v.y };
}

// atomicAdd / atomicMin with relaxed order at device scope, on global memory.
__device__ __forceinline__ uint32_t gatomic_add(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_add(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void gatomic_min(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_fetch_min(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Groups the active lanes of the wave by tile: returns the mask of lanes with
// the same t (the lane's group).  One ballot per distinct tile, no atomics.
__device__ __forceinline__ uint64_t match_tile(uint32_t t)
{
    uint64_t group = 0;
    bool pending = true;
    while (pending) {
        const uint32_t lt = __builtin_amdgcn_readfirstlane(t);
        const uint64_t m = __ballot(t == lt);
        if (t == lt) {
            group = m;
            pending = false;
        }
    }
    return group;
}

// Order-free append for two predicates (p1, p2), plus the block's count of
// p0: every thread with p1 / p2 set gets a slot of its counter, one atomic
// per counter per 256-thread block and only where the predicate holds
// somewhere in the block (two barriers; every thread of the block must call
// it).  The p0 count is stored to part[blockIdx.x], with no atomic: a counter
// that every block updates serializes those updates at its cache line, and
// every block waiting for the returned value paid for that.
__device__ __forceinline__ void block_append(uint32_t *part, bool p0, uint32_t *c1, bool p1, uint32_t *c2, bool p2,
                                             uint32_t *lds, uint32_t &s1, uint32_t &s2)
{
    const uint64_t m0 = __ballot(p0), m1 = __ballot(p1), m2 = __ballot(p2);
    const uint32_t wave = threadIdx.x >> 6;
    if (lane_id() == 0) {
        lds[wave] = (uint32_t)__popcll(m0);
        lds[4 + wave] = (uint32_t)__popcll(m1);
        lds[8 + wave] = (uint32_t)__popcll(m2);
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint32_t *l = lds + 4 * threadIdx.x;
        const uint32_t a = l[0], b = l[1], c = l[2], d = l[3];
        const uint32_t tot = a + b + c + d;
        if (threadIdx.x == 0) {
            stg(part + blockIdx.x, tot);
        } else {
            const uint32_t base = tot ? gatomic_add(threadIdx.x == 1 ? c1 : c2, tot) : 0u;
            l[0] = base;
            l[1] = base + a;
            l[2] = base + a + b;
            l[3] = base + a + b + c;
        }
    }
    __syncthreads();
    s1 = lds[4 + wave] + popc_below(m1);
    s2 = lds[8 + wave] + popc_below(m2);
}

// Sum of a[0, n) by one workgroup of 64 * NW threads (every thread calls
// it; red holds NW words and is reused after the call's second barrier).
template <int NW>
__device__ __forceinline__ uint32_t block_sum(const uint32_t *a, uint32_t n, uint32_t *red)
{
    uint32_t v = 0;
    for (uint32_t i = threadIdx.x; i < n; i += 64u * NW) v += ldg(a + i);
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();  // red may still be read by a previous call
    if (lane_id() == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) t += red[w];
    return t;
}

// Writes the job's counters to its renderer's pinned host block (threads
// 0 .. words - 1 write), after summing what is kept per block or per tile:
// num_alive from the per-setup-block survivor counts and, for single-pass
// binning, total_pairs from the final per-tile cursors (the exact counts).
// Every thread of the workgroup (64 * NW threads) calls it.
template <int NW>
__device__ __forceinline__ void publish_counters(const JobDesc &J, uint32_t *red)
{
    constexpr uint32_t W = sizeof(FrameCounters) / 4;
    static_assert(offsetof(FrameCounters, num_alive) == 0, "num_alive is the first counter word");
    constexpr uint32_t kPairs = offsetof(FrameCounters, total_pairs) / 4;
    constexpr uint32_t kCapTotal = offsetof(FrameCounters, capacity_total) / 4;
    constexpr uint32_t kCapWanted = offsetof(FrameCounters, capacity_wanted) / 4;
    const uint32_t alive = block_sum<NW>(J.alive_part + J.alive_part0, J.alive_parts, red);
    const uint32_t pairs = J.single ? block_sum<NW>(J.tile_count, J.fp.tiles_x * J.fp.tiles_y, red) : 0u;
    if (threadIdx.x < W) {
        uint32_t v = reinterpret_cast<const uint32_t *>(J.counters)[threadIdx.x];
        if (threadIdx.x == 0) v = alive;
        if (J.single) {
            // the queue regions as placed (at most the storage) and as wanted
            const unsigned long long top = J.counters->capacity_top;
            if (threadIdx.x == kPairs) v = pairs;
            if (threadIdx.x == kCapTotal) v = (uint32_t)(top < J.queue_cap ? top : J.queue_cap);
            if (threadIdx.x == kCapWanted) v = (uint32_t)(top < 0xFFFFFFFFull ? top : 0xFFFFFFFFull);
        }
        reinterpret_cast<volatile uint32_t *>(J.host_counters)[threadIdx.x] = v;
    }
}

// counters[t] += 1 for every active lane: one atomic per distinct tile, all
// issued by the group leaders in one instruction; nothing waits for a result.
__device__ __forceinline__ void count_tile(uint32_t *counters, uint32_t t)
{
    const uint64_t g = match_tile(t);
    if (popc_below(g) == 0) gatomic_add(counters + t, (uint32_t)__popcll(g));
}

// Returns a unique slot in tile t's queue (previous cursor value): the group
// leaders reserve their group's slots with one atomic each, in one round trip.
__device__ __forceinline__ uint32_t claim_slot(uint32_t *cursor, uint32_t t)
{
    const uint64_t g = match_tile(t);
    const uint32_t rank = popc_below(g);
    const int leader = __builtin_ctzll(g);
    uint32_t base = 0;
    if (rank == 0) base = gatomic_add(cursor + t, (uint32_t)__popcll(g));
    base = __shfl(base, leader);
    return base + rank;
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
constexpr int kMaskChunks = 64;  // 64-item chunks per start-mask window (one mask word per lane)
constexpr uint32_t kFragRing = 128;  // per-wave queue of covered items (setup's micro triangles, resolve's queue)
constexpr uint32_t kOwnerWords = 64;  // resolve: 64-item mask words per item window

// Depth-test key: min z first, then max ID (low word = ~(id + 1)).
__device__ __forceinline__ uint64_t make_key(float z, uint32_t id)
{
    const uint32_t zb = __float_as_uint(z) & 0x7FFFFFFFu;  // z in [0,1]; -0.0 -> +0.0
    return ((uint64_t)zb << 32) | (uint32_t)~(id + 1u);
}

// Raster-ready triangle: screen vertices, the three edge vectors of
// CCR/CUDAROPs.cu:131-133, depths, |area| and a pixel range.
struct TriLds {
    float sx0, sy0, sx1, sy1, sx2, sy2;
    float ax, ay, bx, by, cx, cy;
    float z0, z1, z2, area;
    uint32_t xr, yr;  // x0 | x1 << 16, y0 | y1 << 16 (absolute pixels)
    uint32_t id;
    float inv;        // queued triangles: 1 / (x1 - x0), for item -> pixel decoding;
                      // screen-list triangles: 1 / area
};

// Setup's micro triangles in LDS: TriLds without the edge vectors, which are
// recomputed (the same binary32 differences) where they are needed.  56 bytes
// instead of 80 leave room for eight waves per SIMD in the setup pass.
struct MicroLds {
    float sx0, sy0, sx1, sy1, sx2, sy2;
    float z0, z1, z2, area;
    uint32_t xr, yr;  // x0 | x1 << 16, y0 | y1 << 16 (absolute pixels)
    uint32_t id;
    float inv;        // 1 / (x1 - x0), for item -> pixel decoding
};

__device__ __forceinline__ void edge_vectors(const TriLds &t, float &ax, float &ay, float &bx, float &by, float &cx,
                                             float &cy)
{
    ax = t.ax; ay = t.ay; bx = t.bx; by = t.by; cx = t.cx; cy = t.cy;
}

// The same differences as set_edges.
__device__ __forceinline__ void edge_vectors(const MicroLds &t, float &ax, float &ay, float &bx, float &by, float &cx,
                                             float &cy)
{
    ax = t.sx1 - t.sx0; ay = t.sy1 - t.sy0;
    bx = t.sx2 - t.sx1; by = t.sy2 - t.sy1;
    cx = t.sx0 - t.sx2; cy = t.sy0 - t.sy2;
}

// Item q of a triangle's pixel range [x0,x1) x [y0,y1), row-major, for
// q < 1024 and a row width w = x1 - x0 <= 1024 (setup's micro triangles: q <
// micro_limit; resolve's items: q < small_limit; both limits are at most
// 1024) with t.inv = fl(1 / w).  (q + 0.5) / w lies at least 0.5 / w >= 2^-11
// from an integer, and fl(fl(q + 0.5) t.inv) is within 2^-23 (q + 0.5) / w <
// 1.3e-4 of it (two roundings; q + 0.5 is exact), so its truncation is the
// row.
template <class Tri>
__device__ __forceinline__ void item_pixel(const Tri &t, uint32_t q, int &px, int &py)
{
    const uint32_t x0 = t.xr & 0xFFFFu, w = (t.xr >> 16) - x0, y0 = t.yr & 0xFFFFu;
    const uint32_t r = (uint32_t)(((float)q + 0.5f) * t.inv);
    px = (int)(x0 + q - __umul24(r, w));
    py = (int)(y0 + r);
}

// The three edge functions of CCR/CUDAROPs.cu:120-145 at pixel (px, py).
template <class Tri>
__device__ __forceinline__ void edges(const Tri &t, float px, float py, float &w0, float &w1, float &w2)
{
    float ax, ay, bx, by, cx, cy;
    edge_vectors(t, ax, ay, bx, by, cx, cy);
    w2 = (px - t.sx0) * ay - (py - t.sy0) * ax;
    w0 = (px - t.sx1) * by - (py - t.sy1) * bx;
    w1 = (px - t.sx2) * cy - (py - t.sy2) * cx;
}

// Depth of RasterizePixelAtomic (CCR/CUDAROPs.cu:329-345) with the D2 window:
// weights /= |area|, z = w0 z0 + w1 z1 + w2 z2, accept 0 <= z <= 1.
template <class Tri>
__device__ __forceinline__ bool depth_of(const Tri &t, float w0, float w1, float w2, float &z)
{
    const float a0 = w0 / t.area, a1 = w1 / t.area, a2 = w2 / t.area;
    const float zi = (a0 * t.z0) + (a1 * t.z1) + (a2 * t.z2);
    if (!(zi >= 0.f && zi <= 1.f)) return false;
    z = zi;
    return true;
}

// Fragment test: all three edge functions >= 0, then the depth.
template <class Tri>
__device__ __forceinline__ bool fragment(const Tri &t, float px, float py, float &z)
{
    float w0, w1, w2;
    edges(t, px, py, w0, w1, w2);
    if (!(w0 >= 0.f && w1 >= 0.f && w2 >= 0.f)) return false;
    return depth_of(t, w0, w1, w2, z);
}

// fragment() at a pixel whose current depth is dbest (<= 1), with a reject
// that needs no division first.  With r = fl(1 / area) (t.inv) and
// u = 2^-24, b_i = fl(w_i r) is within 3.0001 u of a_i = fl(w_i / area), and
// both three-term sums are within 3.0001 u S of their exact values
// (S = sum |a_i z_i|), so z' = fl(fl(b0 z0 + b1 z1) + b2 z2) is within
// 9.0003 u S of the depth that depth_of() computes; S <= S' (1 + 6.01 u) for
// the evaluated S' = sum |fl(b_i z_i)|.  m = 10 u S' + 1e-37 therefore covers
// the difference (the constant covers gradual underflow): where
// z' - m > dbest the fragment loses (or exceeds 1), where z' + m < 0 it is out
// of range.  A NaN or infinite z' or m rejects nothing.
__device__ __forceinline__ bool fragment_closer(const TriLds &t, float px, float py, float dbest, float &z)
{
    float w0, w1, w2;
    edges(t, px, py, w0, w1, w2);
    if (!(w0 >= 0.f && w1 >= 0.f && w2 >= 0.f)) return false;
    const float p0 = (w0 * t.inv) * t.z0, p1 = (w1 * t.inv) * t.z1, p2 = (w2 * t.inv) * t.z2;
    const float zq = (p0 + p1) + p2;
    const float m = ((fabsf(p0) + fabsf(p1)) + fabsf(p2)) * (10.0f / 16777216.0f) + 1.0e-37f;
    if (zq - m > dbest || zq + m < 0.f) return false;
    return depth_of(t, w0, w1, w2, z);
}

__device__ __forceinline__ void set_edges(TriLds &t)
{
    t.ax = t.sx1 - t.sx0; t.ay = t.sy1 - t.sy0;
    t.bx = t.sx2 - t.sx1; t.by = t.sy2 - t.sy1;
    t.cx = t.sx0 - t.sx2; t.cy = t.sy0 - t.sy2;
}

__device__ __forceinline__ void load_tri(const uint32_t *__restrict__ rec, uint32_t id, TriLds &t)
{
    const uint4 *r = reinterpret_cast<const uint4 *>(rec + (size_t)id * kTriRecWords);
    const uint4 q0 = ldg(r), q1 = ldg(r + 1), q2 = ldg(r + 2);
    t.sx0 = __uint_as_float(q0.x); t.sy0 = __uint_as_float(q0.y);
    t.sx1 = __uint_as_float(q0.z); t.sy1 = __uint_as_float(q0.w);
    t.sx2 = __uint_as_float(q1.x); t.sy2 = __uint_as_float(q1.y);
    t.z0 = __uint_as_float(q1.z); t.area = __uint_as_float(q1.w);
    t.z2 = __uint_as_float(q2.x); t.z1 = __uint_as_float(q2.y);
    t.xr = q2.z;  // bbox lo (xmin | ymin << 16)
    t.yr = q2.w;  // bbox hi (xmax | ymax << 16)
    set_edges(t);
    t.id = id;
    t.inv = 0.f;
}

// The depth plane of a screen-list triangle over its pixel range (the
// clipped bbox [x0, x1) x [y0, y1)), in double, formed once per frame by the
// setup pass.  For edge j (vertex (ox, oy), float edge vector (ex, ey)) the
// computed edge value at pixel p is within
//   eps_j = 3.0001 u (Dx |ey| + Dy |ex|)
// of W_j(p) = (px - ox) ey - (py - oy) ex (u = 2^-24; Dx, Dy the largest
// |px - ox|, |py - oy| over the range's pixels): two roundings in each
// product term, one in the difference.  With a_j = fl(w_j / area) and the
// rounded three-term sum, the computed depth is within
//   B = sum_j |z_j| (eps_j + 4.0001 u (|W_j|max + eps_j)) / area
// of the affine L(p) = sum_j z_j W_j(p) / area.  Returns (L(x0, y0), dL/dx,
// dL/dy, M) in float, M >= 1.01 B + 1e-12 |L terms| (the double rounding of
// L), rounded up; M = +inf where no plane applies (non-finite or huge
// values, or M >= 0.2).  Every pixel the per-pixel test admits lies in the
// range, so the bound holds for each tile's share of it.
__device__ __forceinline__ float4 screen_plane(const TriLds &t)
{
    const float4 none = make_float4(0.f, 0.f, 0.f, __int_as_float(0x7f800000));
    const float vmax = fmaxf(fmaxf(fmaxf(fmaxf(fabsf(t.sx0), fabsf(t.sy0)), fmaxf(fabsf(t.sx1), fabsf(t.sy1))),
                                   fmaxf(fmaxf(fabsf(t.sx2), fabsf(t.sy2)), fmaxf(fabsf(t.ax), fabsf(t.ay)))),
                             fmaxf(fmaxf(fabsf(t.bx), fabsf(t.by)), fmaxf(fabsf(t.cx), fabsf(t.cy))));
    if (!(vmax < 1.0e18f)) return none;
    const double A = (double)t.area;
    const double zmax = fmax(fmax(fabs((double)t.z0), fabs((double)t.z1)), fabs((double)t.z2));
    if (!(A > 0.0 && zmax < 1.0e6)) return none;
    // edge j runs from vertex j along (ex, ey); its function weights depth zz[j]
    const double u = 1.0 / 16777216.0;
    const double ox[3] = { t.sx0, t.sx1, t.sx2 }, oy[3] = { t.sy0, t.sy1, t.sy2 };
    const double ex[3] = { t.ax, t.bx, t.cx }, ey[3] = { t.ay, t.by, t.cy };
    const double zz[3] = { t.z2, t.z0, t.z1 };
    const double xa = (double)(t.xr & 0xFFFFu), xb = (double)((t.yr & 0xFFFFu) - 1u);
    const double ya = (double)(t.xr >> 16), yb = (double)((t.yr >> 16) - 1u);
    double L0 = 0.0, Lx = 0.0, Ly = 0.0, B = 0.0, Lmag = 0.0;
    for (int j = 0; j < 3; ++j) {
        const double dxa = xa - ox[j], dxb = xb - ox[j], dya = ya - oy[j], dyb = yb - oy[j];
        const double Dx = fmax(fabs(dxa), fabs(dxb)), Dy = fmax(fabs(dya), fabs(dyb));
        const double w00 = dxa * ey[j] - dya * ex[j], w10 = dxb * ey[j] - dya * ex[j];
        const double w01 = dxa * ey[j] - dyb * ex[j], w11 = dxb * ey[j] - dyb * ex[j];
        const double Wabs = fmax(fmax(fabs(w00), fabs(w10)), fmax(fabs(w01), fabs(w11)));
        const double eps = 3.0001 * u * (Dx * fabs(ey[j]) + Dy * fabs(ex[j]));
        B += fabs(zz[j]) * (eps + 4.0001 * u * (Wabs + eps));
        Lmag += fabs(zz[j]) * Wabs;
        L0 += zz[j] * w00;
        Lx += zz[j] * ey[j];
        Ly -= zz[j] * ex[j];
    }
    const double rA = 1.0 / A;
    const double M = 1.01 * B * rA + 1.0e-12 * Lmag * rA + 1.0e-30;
    if (!(M < 0.2)) return none;
    return make_float4((float)(L0 * rA), (float)(Lx * rA), (float)(Ly * rA), (float)(M * (1.0 + 1.0e-6)));
}

// ---------------------------------------------------------------- kernels

// Block blk of the init stage: resets 256 tiles of its job's per-tile
// counts, min depths and (previous-frame capacities) fallback offsets.
__device__ __forceinline__ void frame_init_block(const JobDesc *__restrict__ jobs, const uint16_t *__restrict__ table,
                                                 uint32_t blk)
{
    uint32_t local;
    const JobDesc &J = jobs[job_of(table, jobs, ST_INIT, local, blk)];
    const uint32_t n = J.fp.tiles_x * J.fp.tiles_y;
    const uint32_t i = local * 256u + threadIdx.x;
    if (i < n) {
        J.tile_count[i] = 0;
        J.tile_min_depth[i] = 1.0f;
        if (J.single && !J.learned) J.tile_pool[i] = 0xFFFFFFFFu;
    }
}

// Texel i of the interleaved copy holds texel i of the diffuse, normal,
// specular and gloss maps (one size), in that order.
__global__ __launch_bounds__(256) void k_interleave_maps(const uint32_t *__restrict__ d, const uint32_t *__restrict__ n,
                                                         const uint32_t *__restrict__ sp, const uint32_t *__restrict__ g,
                                                         uint4 *__restrict__ out, size_t count)
{
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < count; i += (size_t)gridDim.x * 256u)
        out[i] = make_uint4(d[i], n[i], sp[i], g[i]);
}

__global__ __launch_bounds__(256) void k_clear_keys(unsigned long long *__restrict__ keys, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256u) keys[i] = kClearKey;
}

// VS_Kernel / GetNDCVertex (CCR/CUDARenderer.cu:673-700, CCR/CUDAROPs.cu:99-118).
// A block's 256 vertices are 16 KB of input records and 16 KB of attribute
// records per job, contiguous in HBM.  Both go through LDS so that global
// loads and stores run over consecutive 16-byte words (eight full 128-byte
// lines per instruction) instead of one 64-byte record per lane (32 lines
// per instruction, a quarter of each).  Quad q of LDS row r sits at slot
// q ^ ((r >> 2) & 3) of the row: both the row accesses (lane r takes row r)
// and the consecutive ones (lane w takes quad w) then hit distinct banks.
__device__ __forceinline__ void vertex_block(const DevMesh *__restrict__ meshes,
                                             const uint32_t *__restrict__ block_group,
                                             const uint2 *__restrict__ groups,
                                             const uint32_t *__restrict__ entries,
                                             const JobDesc *__restrict__ jobs, uint32_t blk, float4 *rows)
{
    auto at = [](uint32_t r, uint32_t q) { return r * 4u + (q ^ ((r >> 2) & 3u)); };
    // a group is one source mesh drawn by one or more jobs (views of one
    // scene): the vertex is read once and transformed for each of them
    const uint2 gr = groups[block_group[blk]];
    const DevMesh &M0 = meshes[entries[gr.x]];
    static_assert(kVertexChunk <= 256 && 256 % kVertexChunk == 0 && kVertexChunk % 4 == 0,
                  "a chunk's outputs are whole rounds of the block and its input starts on a 16-byte word");
    const uint32_t vbase = (blk - M0.vs_block0) * kVertexChunk;
    const uint32_t nv = M0.num_verts - vbase < kVertexChunk ? M0.num_verts - vbase : kVertexChunk;  // vertices of this block
    const uint32_t tid = threadIdx.x;
    // the block's 56-byte input vertices are one contiguous run of 16-byte
    // words (its start is a multiple of 4 vertices; the last word may reach
    // into the mesh buffer's 16 bytes of padding), copied to LDS as they are
    const float4 *src = reinterpret_cast<const float4 *>(M0.ivert + (size_t)vbase * 14);
    const uint32_t nq = (nv * 14u + 3u) / 4u;
#pragma unroll
    for (int k = 0; k < (int)((kVertexChunk * 14 / 4 + 255) / 256); ++k) {
        const uint32_t w = (uint32_t)k * 256u + tid;
        if (w < nq) rows[w] = ldg(src + w);
    }
    __syncthreads();
    const bool live = tid < nv;
    // the lane's vertex: 7 float2 at float 14 * tid (px py pz u | v nx ny nz | tx ty tz; the colour,
    // floats 11-13, is read at shading)
    float2 f[6] = {};
    if (tid < kVertexChunk) {
        const float2 *vr = reinterpret_cast<const float2 *>(rows) + tid * 7u;
#pragma unroll
        for (int k = 0; k < 6; ++k) f[k] = vr[k];
    }
    const uint32_t v = vbase + tid;
    const float px = f[0].x, py = f[0].y, pz = f[1].x;
    const float u = f[1].y, vv = f[2].x;
    const float nx = f[2].y, ny = f[3].x, nz = f[3].y;
    const float tx = f[4].x, ty = f[4].y, tz = f[5].x;
    if (jobs[M0.job].sattr) {
        // views of one scene (a launch's jobs drawing this mesh share sattr):
        // the view-independent attributes once, then per view the position
        // and one float4 (1/w, view direction times 1/w)
        const JobDesc &J0 = jobs[M0.job];
        const float *Wd = M0.world;
        const float n0 = m4(Wd, 0, 0) * nx + m4(Wd, 0, 1) * ny + m4(Wd, 0, 2) * nz;
        const float n1 = m4(Wd, 1, 0) * nx + m4(Wd, 1, 1) * ny + m4(Wd, 1, 2) * nz;
        const float n2 = m4(Wd, 2, 0) * nx + m4(Wd, 2, 1) * ny + m4(Wd, 2, 2) * nz;
        const float t0 = m4(Wd, 0, 0) * tx + m4(Wd, 0, 1) * ty + m4(Wd, 0, 2) * tz;
        const float t1 = m4(Wd, 1, 0) * tx + m4(Wd, 1, 1) * ty + m4(Wd, 1, 2) * tz;
        const float t2 = m4(Wd, 2, 0) * tx + m4(Wd, 2, 1) * ty + m4(Wd, 2, 2) * tz;
        const float w0 = m4(Wd, 0, 0) * px + m4(Wd, 0, 1) * py + m4(Wd, 0, 2) * pz + m4(Wd, 0, 3);
        const float w1 = m4(Wd, 1, 0) * px + m4(Wd, 1, 1) * py + m4(Wd, 1, 2) * pz + m4(Wd, 1, 3);
        const float w2 = m4(Wd, 2, 0) * px + m4(Wd, 2, 1) * py + m4(Wd, 2, 2) * pz + m4(Wd, 2, 3);
        __syncthreads();  // the input rows are consumed
        rows[at(tid, 0)] = make_float4(u, vv, n0, n1);
        rows[at(tid, 1)] = make_float4(n2, t0, t1, t2);
        __syncthreads();
        float4 *sdst = J0.sattr + (size_t)(M0.vert_base + vbase) * 2;
#pragma unroll
        for (int k = 0; k < (int)(kVertexChunk * 2 / 256); ++k) {
            const uint32_t w = (uint32_t)k * 256u + tid;
            if (w < nv * 2u) stg(sdst + w, rows[at(w >> 1, w & 1u)]);
        }
        for (uint32_t e = 0; e < gr.y; ++e) {
            const DevMesh &M = meshes[entries[gr.x + e]];
            const JobDesc &J = jobs[M.job];
            const float *W = M.wvp;
            float o0 = m4(W, 0, 0) * px + m4(W, 0, 1) * py + m4(W, 0, 2) * pz + m4(W, 0, 3);
            float o1 = m4(W, 1, 0) * px + m4(W, 1, 1) * py + m4(W, 1, 2) * pz + m4(W, 1, 3);
            float o2 = m4(W, 2, 0) * px + m4(W, 2, 1) * py + m4(W, 2, 2) * pz + m4(W, 2, 3);
            const float o3 = m4(W, 3, 0) * px + m4(W, 3, 1) * py + m4(W, 3, 2) * pz + m4(W, 3, 3);
            o0 /= o3;
            o1 /= o3;
            o2 /= o3;
            float vx = w0 - M.cam_pos[0], vy = w1 - M.cam_pos[1], vz = w2 - M.cam_pos[2];
            normalize3(vx, vy, vz);
            const float iw = 1.f / o3;
            if (live) {
                const uint32_t g = M.vert_base + v;
                stg(J.pos4 + g, make_float4(o0, o1, o2, o3));
                stg(J.attr + g, make_float4(iw, vx * iw, vy * iw, vz * iw));
            }
        }
        return;
    }
    for (uint32_t e = 0; e < gr.y; ++e) {
        const DevMesh &M = meshes[entries[gr.x + e]];
        const JobDesc &J = jobs[M.job];
        float4 *__restrict__ pos4 = J.pos4;
        float4 *__restrict__ attr = J.attr;
        const float *W = M.wvp, *Wd = M.world;
        float o0 = m4(W, 0, 0) * px + m4(W, 0, 1) * py + m4(W, 0, 2) * pz + m4(W, 0, 3);
        float o1 = m4(W, 1, 0) * px + m4(W, 1, 1) * py + m4(W, 1, 2) * pz + m4(W, 1, 3);
        float o2 = m4(W, 2, 0) * px + m4(W, 2, 1) * py + m4(W, 2, 2) * pz + m4(W, 2, 3);
        const float o3 = m4(W, 3, 0) * px + m4(W, 3, 1) * py + m4(W, 3, 2) * pz + m4(W, 3, 3);
        o0 /= o3;
        o1 /= o3;
        o2 /= o3;
        const uint32_t g = M.vert_base + v;
        if (live) stg(pos4 + g, make_float4(o0, o1, o2, o3));
        const float n0 = m4(Wd, 0, 0) * nx + m4(Wd, 0, 1) * ny + m4(Wd, 0, 2) * nz;
        const float n1 = m4(Wd, 1, 0) * nx + m4(Wd, 1, 1) * ny + m4(Wd, 1, 2) * nz;
        const float n2 = m4(Wd, 2, 0) * nx + m4(Wd, 2, 1) * ny + m4(Wd, 2, 2) * nz;
        const float t0 = m4(Wd, 0, 0) * tx + m4(Wd, 0, 1) * ty + m4(Wd, 0, 2) * tz;
        const float t1 = m4(Wd, 1, 0) * tx + m4(Wd, 1, 1) * ty + m4(Wd, 1, 2) * tz;
        const float t2 = m4(Wd, 2, 0) * tx + m4(Wd, 2, 1) * ty + m4(Wd, 2, 2) * tz;
        const float w0 = m4(Wd, 0, 0) * px + m4(Wd, 0, 1) * py + m4(Wd, 0, 2) * pz + m4(Wd, 0, 3);
        const float w1 = m4(Wd, 1, 0) * px + m4(Wd, 1, 1) * py + m4(Wd, 1, 2) * pz + m4(Wd, 1, 3);
        const float w2 = m4(Wd, 2, 0) * px + m4(Wd, 2, 1) * py + m4(Wd, 2, 2) * pz + m4(Wd, 2, 3);
        float vx = w0 - M.cam_pos[0], vy = w1 - M.cam_pos[1], vz = w2 - M.cam_pos[2];
        normalize3(vx, vy, vz);
        // Every attribute is stored multiplied by 1/w: the interpolation of
        // CCR/CUDAROPs.cu:347-380 forms w_k * (c_k * (1/w_k)) per pixel, and the
        // product c_k * (1/w_k) is the same binary32 value whether it is formed
        // here once or per pixel.  48 bytes a vertex, the three float4s shading
        // reads (the vertex colour is multiplied there; a 64-byte record with
        // it took the 4K vertex pass 3 us longer, and planar arrays cost shading
        // 3 us more in cache lines touched).
        const float iw = 1.f / o3;
        __syncthreads();  // the rows of the previous job (or of the input) are consumed
        rows[at(tid, 0)] = make_float4(u * iw, vv * iw, iw, n0 * iw);
        rows[at(tid, 1)] = make_float4(n1 * iw, n2 * iw, t0 * iw, t1 * iw);
        rows[at(tid, 2)] = make_float4(t2 * iw, vx * iw, vy * iw, vz * iw);
        __syncthreads();
        float4 *dst = attr + (size_t)(M.vert_base + vbase) * 3;
#pragma unroll
        for (int k = 0; k < (int)((kVertexChunk * 3 + 255) / 256); ++k) {
            const uint32_t w = (uint32_t)k * 256u + tid;
            if (w < nv * 3u) stg(dst + w, rows[at(w / 3u, w % 3u)]);
        }
    }
}

// The vertex launch.  With single-pass jobs (H > 0) its first cap_blocks
// workgroups compute their tiles' capacities and queue regions instead
// (rast_capacity.h): those read only the previous frame's counts and depths,
// so they run beside the vertex transform in the same launch; the LDS is one
// array sized for the larger of the two uses (a capacity model of hidden
// width 64 stages 19 KB, which keeps 8 workgroups per CU, as the vertex
// pass's 16 KB does).  A launch of cap_blocks workgroups alone runs the
// capacity stage by itself (per-stage timing).
template <int H>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_vertex(const DevMesh *__restrict__ meshes,
                                                const uint32_t *__restrict__ block_group,
                                                const uint2 *__restrict__ groups,
                                                const uint32_t *__restrict__ entries,
                                                const JobDesc *__restrict__ jobs,
                                                const uint16_t *__restrict__ cap_table, uint32_t cap_blocks,
                                                const uint16_t *__restrict__ init_table, uint32_t init_blocks)
{
    constexpr int kVertexF4 = 256 * 4, kCapF4 = (capacity_lds_floats<H>() + 3) / 4;
    __shared__ float4 lds[kVertexF4 > kCapF4 ? kVertexF4 : kCapF4];
    if constexpr (H != 0) {
        if (blockIdx.x < cap_blocks) {
            capacity_block<H>(jobs, cap_table, reinterpret_cast<float *>(lds));
            return;
        }
    }
    // the capacity workgroups read the previous frame's counts and depths,
    // the init workgroups reset this frame's (the other parity)
    const uint32_t blk = blockIdx.x - cap_blocks;
    if (blk < init_blocks) {
        frame_init_block(jobs, init_table, blk);
        return;
    }
    vertex_block(meshes, block_group, groups, entries, jobs, blk - init_blocks, lds);
}

// TA_Kernel + TB_Kernel (CCR/CUDARenderer.cu:702-796,903-988) on the CPU
// path's rules (CCR/ERenderer.cpp:73-157, 603-624) with D4/D5, plus the
// micro-triangle raster (RasterizeTriangle, CCR/CUDAROPs.cu:394-407, over the
// whole bbox with the order-free key).
__global__ __launch_bounds__(256) void k_setup(const DevMesh *__restrict__ meshes,
                                               const uint32_t *__restrict__ block_mesh,
                                               const JobDesc *__restrict__ jobs, uint32_t nblocks)
{
    __shared__ uint32_t lds_append[12];
    __shared__ MicroLds mstage[256];
    __shared__ uint32_t moff[256];
    __shared__ unsigned long long mmask[4 * kMaskChunks];
    __shared__ uint32_t mfrag[4 * kFragRing];
    // workgroups in launch order (consecutive blocks spread over the XCDs;
    // XCD-contiguous ranges measured slower)
    const uint32_t blk = blockIdx.x;
    if (blk >= nblocks) return;
    const uint32_t mi = block_mesh[blk];
    const DevMesh &M = meshes[mi];
    const JobDesc &J = jobs[M.job];
    const FrameParams &fp = J.fp;
    const float4 *__restrict__ pos4 = J.pos4;
    uint32_t *__restrict__ trirec = J.trirec;
    uint32_t *__restrict__ tile_count = J.tile_count;
    unsigned long long *__restrict__ keys = J.keys;
    FrameCounters *__restrict__ ctr = J.counters;
    const bool single_pass = J.single != 0;
    const uint32_t k = (blk - M.tri_block0) * 256u + threadIdx.x;
    const bool valid = k < M.num_tris;
    const uint32_t gid = M.tri_base + k;
    bool ok = false;
    int xmin = 0, ymin = 0, xmax = 0, ymax = 0;
    TriLds t;
    if (valid) {
        uint32_t i0, i1, i2;
        if (M.topology == 0) {
            i0 = ldg(M.idx + 3 * (size_t)k); i1 = ldg(M.idx + 3 * (size_t)k + 1); i2 = ldg(M.idx + 3 * (size_t)k + 2);
        } else {  // strip (TA_Kernel, CCR/CUDARenderer.cu:733-746): odd triangles swap their last two indices
            const bool odd = (k & 1u) != 0;
            i0 = ldg(M.idx + k);
            i1 = ldg(M.idx + k + (odd ? 2 : 1));
            i2 = ldg(M.idx + k + (odd ? 1 : 2));
        }
        i0 += M.vert_base; i1 += M.vert_base; i2 += M.vert_base;
        const float4 P0 = ldg(pos4 + i0), P1 = ldg(pos4 + i1), P2 = ldg(pos4 + i2);
        ok = true;
        if (fp.cull_mode != 2) {  // back / front face (CCR/ERenderer.cpp:82-101)
            const float e1x = P1.x - P0.x, e1y = P1.y - P0.y, e1z = P1.z - P0.z;
            const float e2x = P2.x - P0.x, e2y = P2.y - P0.y, e2z = P2.z - P0.z;
            float fx = e1y * e2z - e1z * e2y;
            float fy = e1z * e2x - e1x * e2z;
            float fz = e1x * e2y - e1y * e2x;
            normalize3(fx, fy, fz);
            const float cv = M.cam_fwd[0] * fx + M.cam_fwd[1] * fy + M.cam_fwd[2] * fz;
            ok = fp.cull_mode == 0 ? !(cv <= 0.f) : !(cv >= 0.f);
        }
        // IsTriangleVisible (CCR/CUDAROPs.cu:147-178)
        if (ok && (all_outside(P0.x, P1.x, P2.x, -1.f, 1.f) || all_outside(P0.y, P1.y, P2.y, -1.f, 1.f) ||
                   all_outside(P0.z, P1.z, P2.z, 0.f, 1.f)))
            ok = false;
        if (ok) {
            const float W = (float)fp.width, H = (float)fp.height;
            // NDCToScreenSpace (CCR/CUDAROPs.cu:88-97)
            t.sx0 = ((P0.x + 1) / 2) * W; t.sy0 = ((1 - P0.y) / 2) * H;
            t.sx1 = ((P1.x + 1) / 2) * W; t.sy1 = ((1 - P1.y) / 2) * H;
            t.sx2 = ((P2.x + 1) / 2) * W; t.sy2 = ((1 - P2.y) / 2) * H;
            t.z0 = P0.z; t.z1 = P1.z; t.z2 = P2.z;
            // Cross(v0 - v1, v0 - v2) (CCR/CUDAROPs.cu:336)
            const float ex = t.sx0 - t.sx1, ey = t.sy0 - t.sy1, fx2 = t.sx0 - t.sx2, fy2 = t.sy0 - t.sy2;
            const float cr = ex * fy2 - ey * fx2;
            if (!(cr != 0.f) || !isfinite(cr)) ok = false;
            if (ok) {
                t.area = fabsf(cr);
                // GetBoundingBox with D4 (CCR/CUDAROPs.cu:180-200)
                float mnx = t.sx0, mxx = t.sx0, mny = t.sy0, mxy = t.sy0;
                if (t.sx1 < mnx) mnx = t.sx1;
                if (t.sx2 < mnx) mnx = t.sx2;
                if (t.sx1 > mxx) mxx = t.sx1;
                if (t.sx2 > mxx) mxx = t.sx2;
                if (t.sy1 < mny) mny = t.sy1;
                if (t.sy2 < mny) mny = t.sy2;
                if (t.sy1 > mxy) mxy = t.sy1;
                if (t.sy2 > mxy) mxy = t.sy2;
                xmin = clamp_trunc(mnx, W + 2.f) - 1;
                ymin = clamp_trunc(mny, H + 2.f) - 1;
                xmax = clamp_trunc(mxx, W + 2.f) + 1;
                ymax = clamp_trunc(mxy, H + 2.f) + 1;
                xmin = xmin < 0 ? 0 : xmin;
                ymin = ymin < 0 ? 0 : ymin;
                xmax = xmax > (int)fp.width ? (int)fp.width : xmax;
                ymax = ymax > (int)fp.height ? (int)fp.height : ymax;
                uint4 *r = reinterpret_cast<uint4 *>(trirec + (size_t)gid * kTriRecWords);
                stg(r, make_uint4(__float_as_uint(t.sx0), __float_as_uint(t.sy0), __float_as_uint(t.sx1),
                                  __float_as_uint(t.sy1)));
                stg(r + 1, make_uint4(__float_as_uint(t.sx2), __float_as_uint(t.sy2), __float_as_uint(t.z0),
                                      __float_as_uint(t.area)));
                stg(r + 2, make_uint4(__float_as_uint(t.z2), __float_as_uint(t.z1),
                                      (uint32_t)xmin | ((uint32_t)ymin << 16), (uint32_t)xmax | ((uint32_t)ymax << 16)));
                // shading follows the reference's operation order where the
                // interpolated attributes can be ill-conditioned: a vertex with
                // clip w below 0.75 (at or behind the camera plane, 1/w above
                // 4/3: its premultiplied attributes dominate and cancel), or a
                // mesh whose attributes may exceed 4 in magnitude
                const bool exact = M.shade_exact || !(P0.w >= 0.75f && P1.w >= 0.75f && P2.w >= 0.75f);
                stg(r + 3, make_uint4(i0, i1, i2, mi | (exact ? 0x80000000u : 0u)));
            }
        }
        stg(J.alive + gid, (uint8_t)(ok ? 1 : 0));
    }

    // classification by the pixel range [xmin,xmax) x [ymin,ymax)
    const bool nonempty = ok && xmin < xmax && ymin < ymax;
    const uint32_t npix = nonempty ? (uint32_t)((xmax - xmin) * (ymax - ymin)) : 0u;
    const bool micro = nonempty && npix <= fp.micro_limit;
    uint32_t tx0 = 0, tx1 = 0, ty0 = 0, ty1 = 0, span = 0;
    if (nonempty && !micro) {
        const uint32_t sh = (uint32_t)__builtin_ctz(fp.tile);  // tiles are 8, 16 or 32 pixels
        tx0 = (uint32_t)xmin >> sh; tx1 = (uint32_t)(xmax - 1) >> sh;
        ty0 = (uint32_t)ymin >> sh; ty1 = (uint32_t)(ymax - 1) >> sh;
        span = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
    }
    const bool is_big = nonempty && !micro && span > fp.big_threshold;
    const bool binned = nonempty && !micro && !is_big;

    // Micro triangles: the wave rasterizes the union of their pixel ranges as
    // one item list (prefix sum over the ranges), so a lane's work does not
    // depend on its own triangle's size; fragments go to the key buffer with
    // 64-bit atomic min.  The micro triangles are compacted in LDS; for every
    // 64-item chunk a 64-bit mask marks the items where a triangle starts, so
    // an item's triangle is the count of starts up to it (one popcount).
    // Only about one item in ten is covered, so the edge test runs per item
    // and the covered ones are queued in a per-wave ring in LDS (triangle,
    // offset in its range); the depth (three divisions) and the key minimum
    // run on 64 queued fragments at a time, with every lane busy.
    {
        const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
        const uint64_t mm = __ballot(micro);
        if (mm) {
            const uint32_t rank = popc_below(mm);
            const uint32_t area = micro ? npix : 0u;
            uint32_t incl = area;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o);
                if (lane >= (uint32_t)o) incl += y;
            }
            const uint32_t total = __shfl(incl, 63);
            const uint32_t start = incl - area;
            const uint32_t wsb = wave * 64u;  // the wave's slots in mstage (32-bit LDS index)
            uint32_t *wo = moff + wave * 64;
            unsigned long long *wm = mmask + wave * kMaskChunks;
            if (micro) {
                MicroLds m;
                m.sx0 = t.sx0; m.sy0 = t.sy0; m.sx1 = t.sx1; m.sy1 = t.sy1; m.sx2 = t.sx2; m.sy2 = t.sy2;
                m.z0 = t.z0; m.z1 = t.z1; m.z2 = t.z2; m.area = t.area;
                m.xr = (uint32_t)xmin | ((uint32_t)xmax << 16);
                m.yr = (uint32_t)ymin | ((uint32_t)ymax << 16);
                m.id = gid;
                m.inv = 1.0f / (float)(xmax - xmin);
                mstage[wsb + rank] = m;
                wo[rank] = start;
            }
            uint32_t before = 0;  // triangles starting before the current chunk (wave-uniform)
            uint32_t *ring = mfrag + wave * kFragRing;
            uint32_t head = 0, queued = 0;  // wave-uniform
            // depth test of ring entries [head, head + n): lane l takes entry head + l
            auto drain = [&](uint32_t n) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (lane < n) {
                    const uint32_t e = ring[(head + lane) & (kFragRing - 1u)];
                    const MicroLds &mt = mstage[(uint16_t)(wsb + (e & 63u))];
                    const int px = (int)(mt.xr & 0xFFFFu) + (int)((e >> 6) & 0x3FFu);
                    const int py = (int)(mt.yr & 0xFFFFu) + (int)(e >> 16);
                    float w0, w1, w2, z;
                    edges(mt, (float)px, (float)py, w0, w1, w2);
                    if (depth_of(mt, w0, w1, w2, z))
                        gatomic_min(keys + (size_t)py * fp.width + px, (unsigned long long)make_key(z, mt.id));
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                head += n;
                queued -= n;
            };
            // windows of kMaskChunks chunks (one pass unless micro_limit is large)
            for (uint32_t wbase = 0; wbase < total; wbase += 64u * kMaskChunks) {
                const uint32_t nch = min((total - wbase + 63u) >> 6, (uint32_t)kMaskChunks);
                wm[lane] = 0ull;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (micro && start >= wbase && start - wbase < 64u * kMaskChunks)
                    atomicOr(wm + ((start - wbase) >> 6), 1ull << (start & 63u));
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                for (uint32_t c = 0; c < nch; ++c) {
                    const uint64_t m = wm[c];
                    const uint32_t kk = wbase + c * 64u + lane;
                    const uint32_t j = before + (uint32_t)__popcll(m & ((2ull << lane) - 1ull)) - 1u;
                    bool covered = false;
                    uint32_t entry = 0;
                    if (kk < total) {
                        const MicroLds &mt = mstage[(uint16_t)(wsb + j)];
                        int px, py;
                        item_pixel(mt, kk - wo[j], px, py);
                        float w0, w1, w2;
                        edges(mt, (float)px, (float)py, w0, w1, w2);
                        covered = w0 >= 0.f && w1 >= 0.f && w2 >= 0.f;
                        // offsets in the range: below micro_limit <= 1024 each
                        entry = j | ((uint32_t)(px - (int)(mt.xr & 0xFFFFu)) << 6) |
                                ((uint32_t)(py - (int)(mt.yr & 0xFFFFu)) << 16);
                    }
                    const uint64_t cm = __ballot(covered);
                    if (covered) ring[(head + queued + popc_below(cm)) & (kFragRing - 1u)] = entry;
                    queued += (uint32_t)__popcll(cm);
                    if (queued >= 64u) drain(64u);
                    before += (uint32_t)__popcll(m);
                }
            }
            if (queued) drain(queued);
        }
    }

    // the survivor list feeds the exact binner's fill pass only
    uint32_t bslot, sslot;
    block_append(J.alive_part, ok, &ctr->num_big, is_big, &ctr->num_binned, binned && !single_pass, lds_append, bslot,
                 sslot);
    if (is_big) {
        if (bslot < J.big_cap) stg(J.big + bslot, gid);
        else ctr->storage_overflow = 1;
    }
    if (binned) {
        if (!single_pass) stg(J.survivors + sslot, gid);
        const uint32_t w = tx1 - tx0 + 1;
        uint32_t tx = tx0, ty = ty0;
        for (uint32_t s = 0; s < span; ++s) {
            const uint32_t tl = ty * fp.tiles_x + tx;
            if (single_pass) {
                const uint32_t slot = claim_slot(tile_count, tl);
                const bool fits = slot < ldg(J.tile_cap + tl);
                const uint64_t om = __ballot(!fits);
                if (fits) {
                    const uint32_t pos = ldg(J.tile_off + tl) + slot;
                    if (pos < J.queue_cap) stg(J.queue + pos, gid);
                } else {
                    // capacity exceeded: keep the entry for the exact fallback
                    // (one atomic on the shared counter per wave)
                    const int leader = __builtin_ctzll(om);
                    uint32_t base = 0;
                    if (lane_id() == (uint32_t)leader) base = gatomic_add(&ctr->ovf_entries, (uint32_t)__popcll(om));
                    const uint32_t o = (uint32_t)__shfl((int)base, leader) + popc_below(om);
                    if (o < J.ovf_cap) stg(J.ovf + o, make_uint2(tl, gid));
                    else ctr->storage_overflow = 1;
                }
            } else {
                count_tile(tile_count, tl);
            }
            if (++tx == tx0 + w) { tx = tx0; ++ty; }
        }
    }
    // a screen-list triangle's depth plane over its pixel range, once per
    // frame for every resolve tile (from the record this lane wrote, at the
    // kernel's end, where little else is live)
    if (is_big && bslot < J.big_cap) {
        TriLds bt;
        load_tri(trirec, gid, bt);
        stg(J.big_plane + bslot, screen_plane(bt));
    }
}

// Exclusive scan of a job's exact per-tile counts by one 1024-thread
// workgroup (two-pass jobs): tile_off[n] = the pair count.  Sums are formed
// in 64 bits.  Counts beyond the queue storage set storage_overflow (the
// frame is rendered again with a larger queue); the offsets of such tiles
// are clamped to the storage end.  The fill cursors are zeroed on the way.
__global__ __launch_bounds__(1024) void k_scan(const JobDesc *__restrict__ jobs)
{
    const JobDesc &J = jobs[blockIdx.x];  // one workgroup per job
    if (J.single) return;
    const uint32_t n = J.fp.tiles_x * J.fp.tiles_y;
    const uint32_t *__restrict__ in = J.tile_count;
    uint32_t *__restrict__ out_off = J.tile_off;
    uint32_t *__restrict__ zero_cursor = J.tile_cursor;
    const uint64_t limit = J.queue_cap;
    FrameCounters *__restrict__ ctr = J.counters;
    __shared__ uint64_t wsum[16];
    __shared__ uint64_t carry;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid == 0) carry = 0;
    __syncthreads();
    // chunks of 4096 values: each thread owns 4 consecutive ones (one uint4)
    for (uint32_t base = 0; base < n; base += 4096) {
        const uint32_t i = base + tid * 4;
        uint32_t v[4] = { 0, 0, 0, 0 };
        if (i + 3 < n) {
            const uint4 q = *reinterpret_cast<const uint4 *>(in + i);
            v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
        } else {
            for (uint32_t q = 0; q < 3; ++q)
                if (i + q < n) v[q] = in[i + q];
        }
        const uint64_t s = (uint64_t)v[0] + v[1] + v[2] + v[3];
        uint64_t incl = s;  // inclusive wave prefix
        for (int o = 1; o < 64; o <<= 1) {
            const uint64_t y = __shfl_up(incl, o);
            if (lane >= (uint32_t)o) incl += y;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        uint64_t wbase = carry;
        for (uint32_t w = 0; w < wave; ++w) wbase += wsum[w];
        uint64_t run = wbase + incl - s;
        for (uint32_t q = 0; q < 4; ++q) {
            if (i + q >= n) break;
            out_off[i + q] = (uint32_t)(run < limit ? run : limit);
            run += v[q];
        }
        if (i + 3 < n) *reinterpret_cast<uint4 *>(zero_cursor + i) = make_uint4(0, 0, 0, 0);
        else for (uint32_t q = i; q < n && q < i + 4; ++q) zero_cursor[q] = 0;
        __syncthreads();
        // the chunk's total: the running sum of the last thread that holds values
        if (tid == 1023) carry = wbase + incl;
        __syncthreads();
    }
    if (tid == 0) {
        const uint64_t total = carry;
        out_off[n] = (uint32_t)(total < limit ? total : limit);
        ctr->total_pairs = (uint32_t)(total < 0xFFFFFFFFull ? total : 0xFFFFFFFFull);
        if (total > limit) ctr->storage_overflow = 1;
    }
}

// Exact second pass over the binned triangles (the order inside a queue is
// irrelevant: depth is resolved order-independently).
__global__ __launch_bounds__(256) void k_fill(const JobDesc *__restrict__ jobs, const uint16_t *__restrict__ table)
{
    uint32_t local;
    const JobDesc &J = jobs[job_of(table, jobs, ST_FILL, local)];
    const FrameParams &fp = J.fp;
    const uint32_t *__restrict__ survivors = J.survivors;
    const uint32_t *__restrict__ trirec = J.trirec;
    uint32_t *__restrict__ cursor = J.tile_cursor;
    const uint32_t *__restrict__ tile_off = J.tile_off;
    uint32_t *__restrict__ queue = J.queue;
    const uint32_t queue_cap = J.queue_cap;
    const uint32_t n = J.counters->num_binned;
    const uint32_t stride = J.nblocks[ST_FILL] * 256u;
    for (uint32_t i = local * 256u + threadIdx.x; i < n; i += stride) {
        const uint32_t gid = survivors[i];
        const uint2 bb = *reinterpret_cast<const uint2 *>(trirec + (size_t)gid * kTriRecWords + 10);
        const uint32_t xmin = bb.x & 0xFFFFu, ymin = bb.x >> 16, xmax = bb.y & 0xFFFFu, ymax = bb.y >> 16;
        const uint32_t sh = (uint32_t)__builtin_ctz(fp.tile);  // tiles are 8, 16 or 32 pixels
        const uint32_t tx0 = xmin >> sh, tx1 = (xmax - 1) >> sh;
        const uint32_t ty0 = ymin >> sh, ty1 = (ymax - 1) >> sh;
        for (uint32_t ty = ty0; ty <= ty1; ++ty)
            for (uint32_t tx = tx0; tx <= tx1; ++tx) {
                const uint32_t t = ty * fp.tiles_x + tx;
                const uint32_t pos = tile_off[t] + claim_slot(cursor, t);
                if (pos < queue_cap) queue[pos] = gid;
            }
    }
}

// Single-pass fallback of the previous-frame policy, step 1 (per tile): the
// exact count of a tile is its final fill cursor; the entries beyond its
// capacity get a pool region.  The block allocates for its 256 tiles with one
// atomic and adds its overflowing tiles with one more (no per-wave traffic
// on the shared counters).  Learned capacities overflow so rarely that the
// resolve pass gathers a tile's few entries from the list itself; the
// heuristic's overflow (about fifty times more entries) is placed by these
// two passes, which read the list once in all instead of once per tile.
__global__ __launch_bounds__(256) void k_fallback_alloc(const JobDesc *__restrict__ jobs,
                                                        const uint16_t *__restrict__ table)
{
    __shared__ uint32_t wsum[2][4];
    __shared__ uint32_t base;
    uint32_t local;
    const JobDesc &J = jobs[job_of(table, jobs, ST_FBALLOC, local)];
    const uint32_t n = J.fp.tiles_x * J.fp.tiles_y;
    const uint32_t *__restrict__ count = J.tile_count;
    const uint32_t *__restrict__ cap = J.tile_cap;
    FrameCounters *__restrict__ ctr = J.counters;
    const uint32_t t = local * 256u + threadIdx.x;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    uint32_t excess = 0;
    if (t < n) {
        const uint32_t c = count[t];
        excess = c > cap[t] ? c - cap[t] : 0;
    }
    // inclusive wave prefix of the excess; wave total of overflowed tiles
    uint32_t incl = excess;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o);
        if (lane >= (uint32_t)o) incl += y;
    }
    uint32_t ovt = excess ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) ovt += __shfl_xor(ovt, o);
    if (lane == 63) wsum[0][wave] = incl;
    if (lane == 0) wsum[1][wave] = ovt;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t ex = wsum[0][0] + wsum[0][1] + wsum[0][2] + wsum[0][3];
        const uint32_t tiles = wsum[1][0] + wsum[1][1] + wsum[1][2] + wsum[1][3];
        uint32_t b0 = 0;
        if (ex) {
            b0 = atomicAdd(&ctr->pool_used, ex);
            if (b0 + ex > J.pool_cap || b0 + ex < b0) ctr->storage_overflow = 1;
        }
        base = b0;
        if (tiles) atomicAdd(&ctr->overflow_tiles, tiles);
    }
    __syncthreads();
    if (t < n) {
        uint32_t before = base + incl - excess;
        for (uint32_t q = 0; q < wave; ++q) before += wsum[0][q];
        J.tile_pool[t] = excess ? before : 0xFFFFFFFFu;
        J.tile_cursor[t] = 0;
    }
}

// Step 2 (per overflow entry): exact placement of the entries that did not
// fit.  Entries of one tile come in runs (setup appends them per wave), and
// returned atomics on one address serialize, so each wave takes a tile's
// cursor once per distinct tile: the lanes of that tile get consecutive
// slots from one atomic, and all of the wave's tiles go out in one atomic
// instruction.  The order within a tile's list does not change the
// frame (the depth keys order fragments by depth, then triangle ID).
__global__ __launch_bounds__(256) void k_fallback_fill(const JobDesc *__restrict__ jobs,
                                                       const uint16_t *__restrict__ table)
{
    uint32_t local;
    const JobDesc &J = jobs[job_of(table, jobs, ST_FBFILL, local)];
    const uint2 *__restrict__ ovf = J.ovf;
    const uint32_t *__restrict__ tile_pool = J.tile_pool;
    uint32_t *__restrict__ fb_cursor = J.tile_cursor;
    uint32_t *__restrict__ pool = J.pool;
    const uint32_t pool_cap = J.pool_cap;
    uint32_t n = J.counters->ovf_entries;
    n = n < J.ovf_cap ? n : J.ovf_cap;
    const uint32_t stride = J.nblocks[ST_FBFILL] * 256u;
    const uint32_t lane = lane_id();
    // b is wave-uniform (stride and the wave's offset are multiples of 64)
    for (uint32_t b = local * 256u + (threadIdx.x & ~63u); b < n; b += stride) {
        const uint32_t i = b + lane;
        uint2 e = make_uint2(0u, 0u);
        uint32_t po = 0xFFFFFFFFu;
        if (i < n) {
            e = ldg(ovf + i);
            po = ldg(tile_pool + e.x);
        }
        const bool todo = po != 0xFFFFFFFFu;
        // group the lanes by tile (register work only), then the group
        // leaders take their slots in one atomic instruction
        int leader_of = (int)lane;
        uint32_t rank = 0, group = 0;
        for (uint64_t pend = __ballot(todo); pend;) {
            const int leader = __builtin_ctzll(pend);
            const uint32_t lt = (uint32_t)__builtin_amdgcn_readlane((int)e.x, leader);
            const bool mine = todo && e.x == lt;
            const uint64_t m = __ballot(mine);
            if (mine) {
                leader_of = leader;
                rank = popc_below(m);
            }
            if (lane == (uint32_t)leader) group = (uint32_t)__popcll(m);
            pend &= ~m;
        }
        uint32_t base = 0;
        if (group) base = gatomic_add(fb_cursor + e.x, group);
        base = (uint32_t)__shfl((int)base, leader_of);
        if (todo) {
            const uint32_t pos = po + base + rank;
            if (pos < pool_cap) stg(pool + pos, e.y);
        }
    }
}

// Bounds of a screen-list triangle t over the tile at (X0, Y0), used to skip
// the tile or single pixels before the exact fragment test.  tmax is the
// largest depth the pixels in question hold; depths only decrease from there.
// sp is the triangle's screen_plane.
//
// Tile tests in float, exact.  Every operation of edges() and depth_of() is
// a correctly rounded IEEE operation and rounding is monotone, so a computed
// edge value is monotone in px (py fixed) and in py (px fixed): its extremes
// over the tile's pixels are among the four corner pixels.  An edge negative
// at all four corners fails at every pixel.  The computed depth is monotone
// in each a_i = fl(w_i / area), which lies in [fl(max(lo_i, 0) / area),
// fl(hi_i / area)] at pixels passing all edges; depth_of's sum evaluated
// with each a_i at the matching end bounds the depth of every fragment of the
// tile.  These need finite values below 1e18 in magnitude (no overflow, so
// no NaN); a NaN bound rejects nothing.
//
// Per-pixel plane.  The stored plane, evaluated at the tile corner in float,
// differs from the double plane L by its storage (u of each of L, dL/dx,
// dL/dy: u S at the corner, u span over the tile; S = |L| + |dL/dx dX| +
// |dL/dy dY| with dX, dY the tile corner's offset in the range) and by this
// evaluation (3u S); M adds those, 4u (|L0| + span) for the corner sums
// below and, as the returned plane is evaluated per pixel in float, 8u (|L0|
// + span + 1); the plane is used while M < 0.25.
//
// Returns (x, y, z, w) with lo(p) = x + y (px - X0) + z (py - Y0) <= z(p) <=
// lo(p) + w for every fragment of the tile: the plane (x = L(X0, Y0) - M,
// y = dL/dx, z = dL/dy, w = 2 M + 8u (|L0| + span + 1)) where it applies,
// else the constant float lower bound (w = 1e30).  A fragment loses where
// lo(p) > the pixel's depth and is out of range where lo(p) + w < 0.
// x = +3e38: no pixel of the tile can take a fragment of t (outside an
// edge, depth out of [0, 1], or behind every pixel).
__device__ __forceinline__ float4 tile_bound(const TriLds &t, const float4 sp, int X0, int Y0, int T, float tmax)
{
    const float4 skip = make_float4(3.0e38f, 0.f, 0.f, 0.f);
    const float4 none = make_float4(-1.0e30f, 0.f, 0.f, 1.0e30f);
    const float vmax = fmaxf(fmaxf(fmaxf(fmaxf(fabsf(t.sx0), fabsf(t.sy0)), fmaxf(fabsf(t.sx1), fabsf(t.sy1))),
                                   fmaxf(fmaxf(fabsf(t.sx2), fabsf(t.sy2)), fmaxf(fabsf(t.ax), fabsf(t.ay)))),
                             fmaxf(fmaxf(fabsf(t.bx), fabsf(t.by)), fmaxf(fabsf(t.cx), fabsf(t.cy))));
    if (!(vmax < 1.0e18f)) return none;
    // exact edge extremes over the tile (float, corners)
    float lo0 = 3.0e38f, lo1 = 3.0e38f, lo2 = 3.0e38f, hi0 = -3.0e38f, hi1 = -3.0e38f, hi2 = -3.0e38f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float w0, w1, w2;
        edges(t, (float)(X0 + (c & 1) * (T - 1)), (float)(Y0 + (c >> 1) * (T - 1)), w0, w1, w2);
        lo0 = fminf(lo0, w0); hi0 = fmaxf(hi0, w0);
        lo1 = fminf(lo1, w1); hi1 = fmaxf(hi1, w1);
        lo2 = fminf(lo2, w2); hi2 = fmaxf(hi2, w2);
    }
    if (hi0 < 0.f || hi1 < 0.f || hi2 < 0.f) return skip;
    if (sp.w < 0.2f) {  // the triangle has a plane (+inf: none)
        const float u = 1.0f / 16777216.0f;
        const float dX = (float)(X0 - (int)(t.xr & 0xFFFFu)), dY = (float)(Y0 - (int)(t.xr >> 16));
        const float px = sp.y * dX, py = sp.z * dY;
        const float L0 = (sp.x + px) + py;
        const float S = (fabsf(sp.x) + fabsf(px)) + fabsf(py);
        const float Tm = (float)(T - 1);
        const float span = (fabsf(sp.y) + fabsf(sp.z)) * Tm;
        const float aL = fabsf(L0);
        const float M = (sp.w + u * (4.02f * S + 12.0f * aL + 13.0f * span + 8.0f)) * 1.0001f + 1.0e-30f;
        if (M < 0.25f) {
            // whole-tile tests at the tile's corners (L is affine)
            const float Lmin = (L0 + fminf(sp.y, 0.f) * Tm) + fminf(sp.z, 0.f) * Tm;
            const float Lmax = (L0 + fmaxf(sp.y, 0.f) * Tm) + fmaxf(sp.z, 0.f) * Tm;
            if (Lmin - M > 1.0f || Lmax + M < 0.0f || Lmin - M > tmax) return skip;
            return make_float4(L0 - M, sp.y, sp.z, 2.0f * M + 8.0f * u * (aL + span + 1.0f));
        }
    }
    // no plane: the exact float depth range over the tile
    const float l0 = fmaxf(lo0, 0.f) / t.area, l1 = fmaxf(lo1, 0.f) / t.area, l2 = fmaxf(lo2, 0.f) / t.area;
    const float h0 = hi0 / t.area, h1 = hi1 / t.area, h2 = hi2 / t.area;
    const float zlo = (((t.z0 >= 0.f ? l0 : h0) * t.z0) + ((t.z1 >= 0.f ? l1 : h1) * t.z1)) +
                      ((t.z2 >= 0.f ? l2 : h2) * t.z2);
    const float zhi = (((t.z0 >= 0.f ? h0 : l0) * t.z0) + ((t.z1 >= 0.f ? h1 : l1) * t.z1)) +
                      ((t.z2 >= 0.f ? h2 : l2) * t.z2);
    if (zlo > 1.f || zhi < 0.f || zlo > tmax) return skip;
    return zlo == zlo ? make_float4(fmaxf(zlo, -1.0e30f), 0.f, 0.f, 1.0e30f) : none;
}

// tile_bound's x for "no pixel of the tile": a plane's x stays below 5e5
// (its margin M < 0.25 bounds |x| by 0.25 / (8 u)), a constant bound below 1.
constexpr float kSkipBound = 1.0e37f;

// Float bits that order like the values (negative values reversed).
__device__ __forceinline__ uint32_t ordered_bits(float v)
{
    const uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// The largest depth a tile's pixels hold (block-wide; one barrier).
template <int PPT, int NW>
__device__ __forceinline__ float tile_max_depth(const uint64_t *best, uint32_t lane, uint32_t wave, float *wdmax)
{
    float d = 0.f;
#pragma unroll
    for (int q = 0; q < PPT; ++q) d = fmaxf(d, __uint_as_float((uint32_t)(best[q] >> 32)));
    for (int o = 32; o > 0; o >>= 1) d = fmaxf(d, __shfl_xor(d, o));
    __syncthreads();  // the previous values are read
    if (lane == 0) wdmax[wave] = d;
    __syncthreads();
    float m = wdmax[0];
    for (int w = 1; w < NW; ++w) m = fmaxf(m, wdmax[w]);
    return m;
}

// Ascending bitonic sort of v[0, P) in LDS, P a power of two; ends with a barrier.
template <int NT>
__device__ __forceinline__ void bitonic_sort(uint64_t *v, uint32_t P, uint32_t tid)
{
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < P; i += NT) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint64_t a = v[i], b = v[l];
                    if ((a > b) == ((i & k) == 0)) { v[i] = b; v[l] = a; }
                }
            }
            __syncthreads();
        }
    }
}

// RA_PerTileKernel (CCR/CUDARenderer.cu:1039-1129) redesigned: one workgroup
// per tile, each lane owning PPT pixels.  The lane loads its pixels' keys
// (micro triangles, already rasterized at setup) and resets them for the
// next frame.  The tile's queued triangles (queue region + fallback pool
// region) are rasterized in LDS: a triangle covering at least small_limit
// pixels of the tile one pixel per lane, the others as one compacted item
// list (prefix sum over their clipped pixel ranges, LDS atomic min).  The few
// screen-list triangles are then tested by every lane on its own pixels, with
// a conservative per-pixel depth reject before the exact test.  SHADE != 0
// (1 point, 2 linear sampling): the tile's pixels are shaded right after,
// from the visibility the lanes hold (jobs with that sampling mode and a
// shading stage).
template <bool kLinear, bool kSplit>
__device__ __forceinline__ void shade_pixel(const DevMesh *__restrict__ meshes, const JobDesc &J, const float *unit,
                                            uint32_t x, uint32_t y, uint32_t id);

template <int T, int NT, int STAGE, int SHADE>
__device__ __forceinline__ void resolve_tile(const JobDesc &J, uint32_t tile, const DevMesh *__restrict__ meshes,
                                             const float *unit)
{
    const FrameParams &fp = J.fp;
    const uint32_t *__restrict__ tile_count = J.tile_count;
    const uint32_t *__restrict__ tile_off = J.tile_off;
    const uint32_t *__restrict__ tile_cap = J.single ? J.tile_cap : nullptr;
    const uint32_t *__restrict__ tile_pool = J.tile_pool;
    const uint32_t *__restrict__ queue = J.queue;
    const uint32_t queue_cap = J.queue_cap;
    uint32_t *__restrict__ pool = J.pool;
    const uint32_t pool_cap = J.pool_cap;
    const uint2 *__restrict__ ovf = J.ovf;
    const uint32_t *__restrict__ big = J.big;
    const float4 *__restrict__ big_plane = J.big_plane;
    const uint32_t big_cap = J.big_cap;
    FrameCounters *__restrict__ ctr = J.counters;
    const uint32_t *__restrict__ trirec = J.trirec;
    unsigned long long *__restrict__ keys = J.keys;
    uint32_t *__restrict__ vis = J.vis;
    float *__restrict__ depth = J.depth;
    float *__restrict__ tile_min_depth = J.tile_min_depth;
    const uint32_t small_limit = J.small_limit;
    constexpr int NPIX = T * T;
    constexpr int PPT = NPIX / NT;
    constexpr int NW = NT / 64;
    static_assert(NPIX % NT == 0, "tile pixels must be a multiple of the workgroup size");
    static_assert(T <= 256 && NT <= 256, "ring entries hold the staged triangle and the pixel in 8 bits each");
    __shared__ uint64_t key[NPIX];
    // the queue's staged triangles (56 bytes each: the edge vectors are
    // formed again where they are used, as in the setup pass); once the
    // queue is resolved the same bytes hold the screen-list triangles' depth
    // planes over the tile and their IDs.  Five workgroups fit a CU.
    __shared__ __align__(16) MicroLds stage[NT];
    static_assert(sizeof(stage) >= NT * (sizeof(float4) + sizeof(uint32_t)), "screen-list entries fit the staging area");
    float4 *const bplane = reinterpret_cast<float4 *>(stage);
    uint32_t *const bid = reinterpret_cast<uint32_t *>(bplane + NT);
    __shared__ uint32_t off[NT];
    __shared__ uint32_t wlist[NT];
    __shared__ uint32_t rring[NW * kFragRing];  // per wave: covered items waiting for their depth
    __shared__ unsigned long long omask[kOwnerWords];  // item window: bit k = a staged triangle's first item
    __shared__ uint32_t ocum[kOwnerWords];             // first items before each mask word
    __shared__ uint8_t olist[NT];                      // staged triangles with items, in item order
    __shared__ uint32_t wred_u[NW], wwhole[NW], wown[NW], otot;
    __shared__ float wdmax[NW];
    __shared__ uint32_t ovf_base, ovf_fill;

    const uint32_t ntiles = fp.tiles_x * fp.tiles_y;
    if (tile >= ntiles) return;  // the same for the whole block
    const uint32_t txi = tile % fp.tiles_x, tyi = tile / fp.tiles_x;
    const int X0 = (int)(txi * T), Y0 = (int)(tyi * T);
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;

    // the lane's pixels: p = tid + q * NT (rows of T pixels)
    uint64_t best[PPT];
    int pxs[PPT], pys[PPT];
    bool inside[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int p = (int)tid + q * NT;
        pxs[q] = X0 + p % T;
        pys[q] = Y0 + p / T;
        inside[q] = pxs[q] < (int)fp.width && pys[q] < (int)fp.height;
        best[q] = kClearKey;
        if (inside[q]) {
            const size_t o = (size_t)pys[q] * fp.width + pxs[q];
            best[q] = ldg(keys + o);
            stg(keys + o, (unsigned long long)kClearKey);
        }
    }

    // List A: the tile's queue region.  List B: for single-pass binning,
    // the tile's entries beyond its capacity, which setup appended to the
    // overflow list as (tile, id) pairs.  Learned capacities: the tile's
    // block gathers them itself into a pool region it allocates (one atomic
    // per overflowing tile; the list is read four entries per thread at a
    // time), so no placement pass runs.  Previous-frame capacities: the
    // fallback passes placed them in the tile's pool region.  The order of a
    // tile's entries does not change the frame (the depth keys order
    // fragments by depth, then ID).
    const uint32_t cnt = ldg(tile_count + tile);
    uint32_t n_a = cnt, n_b = 0;
    const uint32_t qoff = ldg(tile_off + tile);
    const uint32_t *list_a = queue + qoff;
    const uint32_t *list_b = pool;
    if (tile_cap) {
        const uint32_t cp = ldg(tile_cap + tile);
        if (cnt > cp && !J.learned) {
            n_a = cp;
            const uint32_t po = ldg(tile_pool + tile);
            if (po != 0xFFFFFFFFu) {
                list_b = pool + po;
                const uint32_t room_b = po < pool_cap ? pool_cap - po : 0;
                n_b = cnt - cp;
                n_b = n_b < room_b ? n_b : room_b;
            }
        } else if (cnt > cp) {  // the same for the whole block
            n_a = cp;
            const uint32_t want = cnt - cp;
            if (tid == 0) {
                const uint32_t b = gatomic_add(&ctr->pool_used, want);
                if (b + want > pool_cap || b + want < b) ctr->storage_overflow = 1;
                gatomic_add(&ctr->overflow_tiles, 1u);
                ovf_base = b;
                ovf_fill = 0;
            }
            __syncthreads();
            const uint32_t pbase = ovf_base;
            uint32_t ne = ldg(&ctr->ovf_entries);
            ne = ne < J.ovf_cap ? ne : J.ovf_cap;
            for (uint32_t b0 = 0; b0 < ne; b0 += 4u * NT) {
                uint2 e[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t i = b0 + (uint32_t)q * NT + tid;
                    e[q] = i < ne ? ldg(ovf + i) : make_uint2(0xFFFFFFFFu, 0u);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const bool mine = e[q].x == tile;
                    const uint64_t m = __ballot(mine);
                    if (m) {
                        const int leader = __builtin_ctzll(m);
                        uint32_t wb = 0;
                        if (lane == (uint32_t)leader) wb = atomicAdd(&ovf_fill, (uint32_t)__popcll(m));
                        wb = (uint32_t)__shfl((int)wb, leader);
                        const uint32_t pos = pbase + wb + popc_below(m);
                        if (mine && pos < pool_cap) stg(pool + pos, e[q].y);
                    }
                }
            }
            __syncthreads();  // the gathered entries are visible to the whole block
            const uint32_t got = ovf_fill;  // = want unless the overflow list was cut short (storage overflow)
            if (pbase < pool_cap) {
                list_b = pool + pbase;
                const uint32_t room_b = pool_cap - pbase;
                n_b = got < room_b ? got : room_b;
            }
        }
    }
    const uint32_t room_a = qoff < queue_cap ? queue_cap - qoff : 0;
    n_a = n_a < room_a ? n_a : room_a;
    const uint32_t n_tile = n_a + n_b;
    const uint32_t num_tris = fp.num_tris;

    if (n_tile) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) key[(int)tid + q * NT] = best[q];
        for (uint32_t base = 0; base < n_tile; base += NT) {
            const uint32_t i = base + tid;
            uint32_t area = 0;
            bool whole = false;
            TriLds t;
            t.xr = 0; t.yr = 0; t.inv = 0.f;
            if (i < n_tile) {
                const uint32_t id = i < n_a ? ldg(list_a + i) : ldg(list_b + (i - n_a));
                if (id < num_tris) {
                    load_tri(trirec, id, t);
                    const int bx0 = (int)(t.xr & 0xFFFFu), by0 = (int)(t.xr >> 16);
                    const int bx1 = (int)(t.yr & 0xFFFFu), by1 = (int)(t.yr >> 16);
                    const int x0 = bx0 > X0 ? bx0 : X0, x1 = bx1 < X0 + T ? bx1 : X0 + T;
                    const int y0 = by0 > Y0 ? by0 : Y0, y1 = by1 < Y0 + T ? by1 : Y0 + T;
                    if (x0 < x1 && y0 < y1) {
                        t.xr = (uint32_t)x0 | ((uint32_t)x1 << 16);
                        t.yr = (uint32_t)y0 | ((uint32_t)y1 << 16);
                        area = (uint32_t)((x1 - x0) * (y1 - y0));
                        if (area >= small_limit) { whole = true; area = 0; }
                        t.inv = 1.0f / (float)(x1 - x0);
                    }
                }
            }
            {
                MicroLds m;
                m.sx0 = t.sx0; m.sy0 = t.sy0; m.sx1 = t.sx1; m.sy1 = t.sy1; m.sx2 = t.sx2; m.sy2 = t.sy2;
                m.z0 = t.z0; m.z1 = t.z1; m.z2 = t.z2; m.area = t.area;
                m.xr = t.xr; m.yr = t.yr; m.id = t.id; m.inv = t.inv;
                stage[tid] = m;
            }
            // block-wide exclusive scan of the item counts
            uint32_t incl = area;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o);
                if (lane >= (uint32_t)o) incl += y;
            }
            if (lane == 63) wred_u[wave] = incl;
            const uint64_t wm = __ballot(whole), hm = __ballot(area != 0);
            if (lane == 0) {
                wwhole[wave] = (uint32_t)__popcll(wm);
                wown[wave] = (uint32_t)__popcll(hm);
            }
            __syncthreads();
            uint32_t before = 0, nitems = 0, pbefore = 0, nwhole = 0, obefore = 0;
            for (int w = 0; w < NW; ++w) {
                if (w < (int)wave) { before += wred_u[w]; pbefore += wwhole[w]; obefore += wown[w]; }
                nitems += wred_u[w];
                nwhole += wwhole[w];
            }
            const uint32_t first = before + incl - area;  // this triangle's first item
            off[tid] = first;
            if (whole) wlist[pbefore + popc_below(wm)] = tid;
            if (area) olist[obefore + popc_below(hm)] = (uint8_t)tid;
            __syncthreads();
            // compacted items, in windows of 64 kOwnerWords items: a mask bit
            // marks each triangle's first item, so an item's triangle is the
            // count of first items up to it (a popcount and the word's prefix
            // count).  The edge test runs per item; the covered items (about
            // one in six) queue in the wave's ring (staged triangle, pixel in
            // the tile), and the depth (three divisions) and key minimum run
            // on 64 of them at a time, with every lane busy.
            {
                uint32_t *ring = rring + wave * kFragRing;
                uint32_t head = 0, queued = 0;  // wave-uniform
                auto drain = [&](uint32_t n) {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (lane < n) {
                        const uint32_t e = ring[(head + lane) & (kFragRing - 1u)];
                        const MicroLds &mt = stage[(uint16_t)(e & 0xFFu)];
                        const int lx = (int)((e >> 8) & 0xFFu), ly = (int)(e >> 16);
                        float w0, w1, w2, z;
                        edges(mt, (float)(X0 + lx), (float)(Y0 + ly), w0, w1, w2);
                        if (depth_of(mt, w0, w1, w2, z))
                            atomicMin((unsigned long long *)&key[ly * T + lx], (unsigned long long)make_key(z, mt.id));
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    head += n;
                    queued -= n;
                };
                constexpr uint32_t kWindow = 64u * kOwnerWords;
                static_assert(kWindow % NT == 0, "a window holds whole rounds of the block's item chunks");
                uint32_t rbase = 0;  // triangles whose first item precedes the window
                for (uint32_t wb = 0; wb < nitems; wb += kWindow) {
                    if (wb) __syncthreads();  // the previous window's masks are read
                    if (tid < kOwnerWords) omask[tid] = 0ull;
                    __syncthreads();
                    if (area && first >= wb && first - wb < kWindow)
                        atomicOr(&omask[(first - wb) >> 6], 1ull << (first & 63u));
                    __syncthreads();
                    if (wave == 0) {
                        const uint32_t c = lane < kOwnerWords ? (uint32_t)__popcll(omask[lane]) : 0u;
                        uint32_t ci = c;
                        for (int o = 1; o < 64; o <<= 1) {
                            const uint32_t y = __shfl_up(ci, o);
                            if (lane >= (uint32_t)o) ci += y;
                        }
                        if (lane < kOwnerWords) ocum[lane] = rbase + ci - c;
                        if (lane == 63) otot = ci;
                    }
                    __syncthreads();
                    rbase += otot;
                    const uint32_t wend = nitems - wb < kWindow ? nitems : wb + kWindow;
                    for (uint32_t kb = wb + wave * 64u; kb < wend; kb += NT) {
                        const uint32_t kk = kb + lane;
                        const uint32_t wi = (kb - wb) >> 6;  // wave-uniform
                        const uint64_t m = omask[wi];
                        bool covered = false;
                        uint32_t entry = 0;
                        if (kk < wend) {
                            const uint32_t j = olist[ocum[wi] + (uint32_t)__popcll(m & ((2ull << lane) - 1ull)) - 1u];
                            const MicroLds &mt = stage[(uint16_t)j];
                            int px, py;
                            item_pixel(mt, kk - off[j], px, py);
                            float w0, w1, w2;
                            edges(mt, (float)px, (float)py, w0, w1, w2);
                            covered = w0 >= 0.f && w1 >= 0.f && w2 >= 0.f;
                            entry = j | ((uint32_t)(px - X0) << 8) | ((uint32_t)(py - Y0) << 16);
                        }
                        const uint64_t cm = __ballot(covered);
                        if (covered) ring[(head + queued + popc_below(cm)) & (kFragRing - 1u)] = entry;
                        queued += (uint32_t)__popcll(cm);
                        if (queued >= 64u) drain(64u);
                    }
                }
                if (queued) drain(queued);
            }
            if (nwhole) {
                __syncthreads();
#pragma unroll
                for (int q = 0; q < PPT; ++q) best[q] = key[(int)tid + q * NT];
                // wave-uniform row range of this wave's pixels (PPT == 1)
                const int wy0 = Y0 + (int)(wave * 64) / T;
                const int wy1 = Y0 + ((int)(wave * 64) + 63) / T + 1;
                for (uint32_t q2 = 0; q2 < nwhole; ++q2) {
                    const MicroLds &st = stage[wlist[q2]];
                    const int x0 = (int)(st.xr & 0xFFFFu), x1 = (int)(st.xr >> 16);
                    const int y0 = (int)(st.yr & 0xFFFFu), y1 = (int)(st.yr >> 16);
                    if (PPT == 1 && (y1 <= wy0 || y0 >= wy1)) continue;
#pragma unroll
                    for (int q = 0; q < PPT; ++q) {
                        const int px = pxs[q], py = pys[q];
                        if (px >= x0 && px < x1 && py >= y0 && py < y1) {
                            float z;
                            if (fragment(st, (float)px, (float)py, z)) {
                                const uint64_t kk = make_key(z, st.id);
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
#pragma unroll
        for (int q = 0; q < PPT; ++q) best[q] = key[(int)tid + q * NT];
    }

    // Screen-list triangles.  A triangle's plane and pixel range are tested
    // first, the whole triangle only where some lane passes them.  The
    // wave's largest depth (pixels of the frame) skips a triangle for the
    // whole wave where its plane lies above it over the entire tile: the
    // plane's float value at its lowest corner, less 8 u S (S its term
    // magnitudes; its float evaluation is within 3 u S of the exact affine
    // plane, whose minimum over the tile is at that corner), bounds the
    // triangle's depth at every pixel from below.  The wave keeps a mask of
    // up to 64 triangles (one per lane) whose bound is not above its largest
    // depth and narrows it each time that depth falls, so a skipped triangle
    // costs no iteration.
    //
    // LDS holds a kept triangle's plane and ID; its record is read again at
    // its (uniform) address where it is run.  A list of at most 64 is
    // handled by every wave on its own: its lanes load and bound the whole
    // list against the wave's largest depth, with no barrier.  A longer list is
    // first bounded over the tile (pass 1), keeping the triangles that can
    // still take a pixel, keyed by their depth lower bound (in the key
    // array, free once the queue is resolved).  Sorted front to back, pass 2
    // then stages them chunk by chunk and stops where the next lower bound
    // lies behind every pixel of the tile.  More survivors than fit go to
    // pass 2 in their stored order.  Each chunk is loaded and bounded in
    // parallel and compacted into LDS.
    uint32_t n_big = ldg(&ctr->num_big);
    n_big = n_big < big_cap ? n_big : big_cap;
    if (n_big) {
        auto wave_max_depth = [&]() {
            float d = 0.f;
#pragma unroll
            for (int q = 0; q < PPT; ++q)
                if (inside[q]) d = fmaxf(d, __uint_as_float((uint32_t)(best[q] >> 32)));
            for (int o = 32; o > 0; o >>= 1) d = fmaxf(d, __shfl_xor(d, o));
            return d;
        };
        auto wave_floor = [](const float4 &pl) {
            const float lmin = pl.x + fminf(pl.y, 0.f) * (float)(T - 1) + fminf(pl.z, 0.f) * (float)(T - 1);
            const float mag = fabsf(pl.x) + (fabsf(pl.y) + fabsf(pl.z)) * (float)(T - 1);
            return lmin - mag * (8.0f / 16777216.0f);
        };
        float wmax = wave_max_depth();
        // the exact tests of triangle st on the lanes' pixels; true when some
        // lane passed the per-pixel bounds (wmax is then current)
        auto run = [&](const TriLds &st, const float4 &pl, const uint2 &bb) {
            const int x0 = (int)(bb.x & 0xFFFFu), y0 = (int)(bb.x >> 16);
            const int x1 = (int)(bb.y & 0xFFFFu), y1 = (int)(bb.y >> 16);
            bool pass[PPT];
            bool any = false;
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int px = pxs[q], py = pys[q];
                // lower bound of the triangle's depth here: above the current
                // depth it loses, and with the band width below 0 it is out of range
                const float zc = pl.x + pl.y * (float)(px - X0) + pl.z * (float)(py - Y0);
                const float db = __uint_as_float((uint32_t)(best[q] >> 32));
                pass[q] = px >= x0 && px < x1 && py >= y0 && py < y1 && !(zc > db) && !(zc + pl.w < 0.f);
                any = any || pass[q];
            }
            if (__ballot(any) == 0) return false;
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                float z;
                if (pass[q] && fragment_closer(st, (float)pxs[q], (float)pys[q],
                                               __uint_as_float((uint32_t)(best[q] >> 32)), z)) {
                    const uint64_t kk = make_key(z, st.id);
                    best[q] = kk < best[q] ? kk : best[q];
                }
            }
            wmax = wave_max_depth();
            return true;
        };
        // the short form: every wave bounds the whole list (at most 64) on
        // its own against its largest depth and keeps it in its own 64
        // entries of bplane / bid
        const bool own = n_big <= 64u;  // the same for the whole block
        float tmax = 1.0f;
        uint64_t *cand = key;  // NPIX entries: (ordered lower bound << 32) | list index
        uint32_t n_cand = 0;
        const bool presort = n_big > (uint32_t)NT;
        if (!own) tmax = tile_max_depth<PPT, NW>(best, lane, wave, wdmax);
        for (uint32_t base = 0; presort && base < n_big; base += NT) {
            const uint32_t i = base + tid;
            bool keep = false;
            float lb = 0.f;
            if (i < n_big) {
                const uint32_t id = ldg(big + i);
                if (id < num_tris) {
                    TriLds t;
                    load_tri(trirec, id, t);
                    const int x0 = (int)(t.xr & 0xFFFFu), y0 = (int)(t.xr >> 16);
                    const int x1 = (int)(t.yr & 0xFFFFu), y1 = (int)(t.yr >> 16);
                    if (x1 > X0 && x0 < X0 + T && y1 > Y0 && y0 < Y0 + T) {
                        const float4 pl = tile_bound(t, ldg(big_plane + i), X0, Y0, T, tmax);
                        keep = !(pl.x > kSkipBound);
                        lb = pl.x + fminf(pl.y, 0.f) * (float)(T - 1) + fminf(pl.z, 0.f) * (float)(T - 1);
                    }
                }
            }
            const uint64_t km = __ballot(keep);
            if (lane == 0) wred_u[wave] = (uint32_t)__popcll(km);
            __syncthreads();
            uint32_t before = 0, nk = 0;
            for (int w = 0; w < NW; ++w) {
                if (w < (int)wave) before += wred_u[w];
                nk += wred_u[w];
            }
            const uint32_t slot = n_cand + before + popc_below(km);
            if (keep && slot < (uint32_t)NPIX) cand[slot] = ((uint64_t)ordered_bits(lb) << 32) | i;
            n_cand += nk;
            __syncthreads();
        }
        const bool sorted = presort && n_cand <= (uint32_t)NPIX;
        if (sorted) {
            uint32_t P = 1;
            while (P < n_cand) P <<= 1;
            for (uint32_t q = n_cand + tid; q < P; q += NT) cand[q] = ~0ull;
            __syncthreads();
            bitonic_sort<NT>(cand, P, tid);
        }
        const uint32_t n_list = sorted ? n_cand : n_big;
        const uint32_t step = own ? 64u : (uint32_t)NT;
        for (uint32_t base = 0; base < n_list; base += step) {
            if (!own) {
                if (base) tmax = tile_max_depth<PPT, NW>(best, lane, wave, wdmax);
                // sorted: every further triangle is bounded below by this one's bound
                if (sorted && (uint32_t)(cand[base] >> 32) > ordered_bits(tmax)) break;
            }
            const uint32_t i = base + (own ? lane : tid);
            bool keep = false;
            uint32_t id = 0xFFFFFFFFu;
            float4 pl = make_float4(3.0e38f, 0.f, 0.f, 0.f);
            if (i < n_list) {
                const uint32_t li = sorted ? (uint32_t)cand[i] : i;
                id = ldg(big + li);
                if (id < num_tris) {
                    TriLds t;
                    load_tri(trirec, id, t);
                    const int x0 = (int)(t.xr & 0xFFFFu), y0 = (int)(t.xr >> 16);
                    const int x1 = (int)(t.yr & 0xFFFFu), y1 = (int)(t.yr >> 16);
                    if (x1 > X0 && x0 < X0 + T && y1 > Y0 && y0 < Y0 + T) {
                        pl = tile_bound(t, ldg(big_plane + li), X0, Y0, T, own ? wmax : tmax);
                        keep = !(pl.x > kSkipBound);
                    }
                }
            }
            // the entries to run: the wave's own 64, or the block's compacted chunk
            uint32_t j0, nb;
            if (own) {
                j0 = wave * 64u;
                nb = n_list;
                if (!keep) pl = make_float4(3.0e38f, 0.f, 0.f, 0.f);
                bplane[j0 + lane] = pl;
                bid[j0 + lane] = id;
                // the wave's lanes read one another's entries below
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            } else {
                const uint64_t km = __ballot(keep);
                if (lane == 0) wred_u[wave] = (uint32_t)__popcll(km);
                __syncthreads();
                uint32_t before = 0;
                nb = 0;
                for (int w = 0; w < NW; ++w) {
                    if (w < (int)wave) before += wred_u[w];
                    nb += wred_u[w];
                }
                if (keep) {
                    const uint32_t slot = before + popc_below(km);
                    bplane[slot] = pl;
                    bid[slot] = id;
                }
                __syncthreads();
                j0 = 0;
            }
            // 64 entries at a time, one per lane
            for (uint32_t g = 0; g < nb; g += 64u) {
                const uint32_t jl = g + lane;
                float floor_l = 3.0e38f;
                if (jl < nb) floor_l = wave_floor(bplane[j0 + jl]);
                uint64_t live = __ballot(jl < nb && !(floor_l > wmax));
                while (live) {
                    const uint32_t j = j0 + g + (uint32_t)__builtin_ctzll(live);
                    live &= live - 1u;
                    const float4 plj = bplane[j];
                    TriLds st;
                    load_tri(trirec, bid[j], st);
                    st.inv = 1.0f / st.area;
                    if (run(st, plj, make_uint2(st.xr, st.yr))) live &= __ballot(!(floor_l > wmax));
                }
            }
            if (!own) __syncthreads();
        }
    }

    // write out and the tile's min depth (a feature of the capacity model)
    float dmin = 1.0f;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const float d = __uint_as_float((uint32_t)(best[q] >> 32));
        dmin = d < dmin ? d : dmin;
        if (inside[q]) {
            const size_t o = (size_t)pys[q] * fp.width + pxs[q];
            stg(vis + o, ~(uint32_t)best[q] - 1u);
            stg(depth + o, d);
        }
    }
    if (tile_min_depth) {
        for (int o = 32; o > 0; o >>= 1) {
            const float y = __shfl_xor(dmin, o);
            dmin = y < dmin ? y : dmin;
        }
        // one block per tile: the block's minimum is the tile's, a plain store
        __syncthreads();  // wdmax may still be read by the screen-list bound
        if (lane == 0) wdmax[wave] = dmin;
        __syncthreads();
        if (tid == 0) {
            float m = wdmax[0];
            for (int w = 1; w < NW; ++w) m = wdmax[w] < m ? wdmax[w] : m;
            stg(tile_min_depth + tile, m);
        }
    }
    if constexpr (SHADE != 0) {
        if (J.nblocks[ST_SHADE] && (fp.sample_state != 0) == (SHADE == 2)) {
#pragma unroll
            for (int q = 0; q < PPT; ++q)
                if (inside[q]) shade_pixel<SHADE == 2, false>(meshes, J, unit, (uint32_t)pxs[q], (uint32_t)pys[q],
                                                       ~(uint32_t)best[q] - 1u);
        }
    }
}

// One workgroup per tile, in launch order: the hardware spreads consecutive
// workgroups over the 8 XCDs, so neighbouring tiles run on different XCDs at
// once (giving each XCD a contiguous band instead measured 30% slower).
template <int T, int NT, int STAGE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5))) void k_resolve(
    const JobDesc *__restrict__ jobs, const uint16_t *__restrict__ table)
{
    uint32_t local;
    const JobDesc &J = jobs[job_of(table, jobs, STAGE, local)];
    resolve_tile<T, NT, STAGE, 0>(J, local, nullptr, nullptr);
}

// ------------------------------------------------------------- shading

// (u8)(c * 255) saturating: 0 for negatives and NaN (fmaxf returns the
// number), 255 from 255 up, truncation between.
__device__ __forceinline__ uint32_t pack_channel(float c)
{
    return (uint32_t)__builtin_fminf(__builtin_fmaxf(c * 255.f, 0.f), 255.f);
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

// Texel indices of CUDATextureSampler::SamplePoint / SampleLinear
// (CCR/CUDATextureSampler.cuh:86-190): the centre texel and, for the linear
// mode, its left, right, lower and upper neighbours (the centre where one
// falls outside the texture).  ok = false outside [0, 1] (magenta); the
// indices are then 0, so the loads stay in bounds and go out with the others.
struct Taps {
    int c, n[4];
    bool ok;
};

__device__ __forceinline__ Taps texel_taps(const DevTexture &t, float u, float v)
{
    Taps k;
    int x = 0, y = 0;
    const bool okx = texel_coord(u, t.w, x), oky = texel_coord(v, t.h, y);
    k.ok = okx && oky;
    const int w = (int)t.w, size = w * (int)t.h;
    k.c = k.ok ? x + y * w : 0;
    k.n[0] = k.c - 1 < 0 ? k.c : k.c - 1;
    k.n[1] = k.c + 1 >= size ? k.c : k.c + 1;
    k.n[2] = k.c + w >= size ? k.c : k.c + w;
    k.n[3] = k.c - w < 0 ? k.c : k.c - w;
    return k;
}

// Colour of point-sampled texel s; unit[c] = c / 255.f (the same correctly
// rounded quotient as the reference).
__device__ __forceinline__ void decode_point(uint32_t s, bool ok, const float *unit, float &r, float &g, float &b)
{
    const float ur = unit[s & 0xFFu], ug = unit[(s >> 8) & 0xFFu], ub = unit[(s >> 16) & 0xFFu];
    r = ok ? ur : 1.f;
    g = ok ? ug : 0.f;
    b = ok ? ub : 1.f;
}

// SampleLinear's weighting: neighbours x 1/8 each in order, then the centre
// / 2, then x (1/255).  Every partial sum of that float accumulation is a
// multiple of 1/8 below 2^11, hence exact, so the sum is (n0 + n1 + n2 + n3 +
// 4 centre) / 8; it is formed here in integers, red and blue side by side in
// 16-bit halves.  S / 8 x (1/255) and S x ((1/255) / 8) round the same real
// number, so the colours are bitwise those of the float accumulation.
__device__ __forceinline__ void decode_linear(const uint32_t *s, uint32_t sc, bool ok, float &r, float &g, float &b)
{
    constexpr uint32_t kRB = 0x00FF00FFu;
    const uint32_t rb = (s[0] & kRB) + (s[1] & kRB) + (s[2] & kRB) + (s[3] & kRB) + ((sc & kRB) << 2);
    const uint32_t gs = ((s[0] >> 8) & 0xFFu) + ((s[1] >> 8) & 0xFFu) + ((s[2] >> 8) & 0xFFu) +
                        ((s[3] >> 8) & 0xFFu) + (((sc >> 8) & 0xFFu) << 2);
    const float k = (1.0f / 255.f) * 0.125f;
    r = ok ? (float)(rb & 0xFFFFu) * k : 1.f;
    g = ok ? (float)gs * k : 0.f;
    b = ok ? (float)(rb >> 16) * k : 1.f;
}

// Texel at byte offset off from base: a 32-bit offset (textures hold at most
// 2^30 texels), so a wave-uniform base stays in scalar registers.
__device__ __forceinline__ uint32_t texel_at(const uint32_t *base, uint32_t off)
{
    return ldg(reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(base) + off));
}

// Colours of the four maps at (u, v): diffuse, normal and specular RGB and
// the gloss map's red channel.
struct MapColours {
    float d[3], n[3], s[3], g;
};

// Samples the four maps.  kShared: all four have the same size, so they
// share the texel indices (computed once).  Point mode decodes through the
// unit table (c / 255, correctly rounded).
template <bool kShared>
__device__ __forceinline__ MapColours sample_maps(const DevTexture (&tex)[4], int mode, float u, float v,
                                                  const float *unit)
{
    MapColours o;
    float gg, gb;
    Taps tp[4];
    tp[0] = texel_taps(tex[0], u, v);
#pragma unroll
    for (int k = 1; k < 4; ++k) tp[k] = kShared ? tp[0] : texel_taps(tex[k], u, v);
    if (mode == 0) {
        uint32_t sc[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) sc[k] = texel_at(tex[k].texels, (uint32_t)tp[k].c * 4u);
        decode_point(sc[0], tp[0].ok, unit, o.d[0], o.d[1], o.d[2]);
        decode_point(sc[1], tp[1].ok, unit, o.n[0], o.n[1], o.n[2]);
        decode_point(sc[2], tp[2].ok, unit, o.s[0], o.s[1], o.s[2]);
        decode_point(sc[3], tp[3].ok, unit, o.g, gg, gb);
    } else {
        uint32_t sc[4], sn[4][4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            sc[k] = texel_at(tex[k].texels, (uint32_t)tp[k].c * 4u);
#pragma unroll
            for (int i = 0; i < 4; ++i) sn[k][i] = texel_at(tex[k].texels, (uint32_t)tp[k].n[i] * 4u);
        }
        decode_linear(sn[0], sc[0], tp[0].ok, o.d[0], o.d[1], o.d[2]);
        decode_linear(sn[1], sc[1], tp[1].ok, o.n[0], o.n[1], o.n[2]);
        decode_linear(sn[2], sc[2], tp[2].ok, o.s[0], o.s[1], o.s[2]);
        decode_linear(sn[3], sc[3], tp[3].ok, o.g, gg, gb);
    }
    (void)gg; (void)gb;
    return o;
}

// sample_maps for the interleaved copy of four same-size maps (texel
// indices from the diffuse map's size): one 16-byte load per tap instead of
// one 4-byte load per tap and map; the same texels, decoded the same way.
__device__ __forceinline__ MapColours sample_quads(const uint4 *texq, const DevTexture &t0, int mode, float u, float v,
                                                   const float *unit)
{
    MapColours o;
    float gg, gb;
    const Taps tp = texel_taps(t0, u, v);
    const char *base = reinterpret_cast<const char *>(texq);
    auto quad = [base](int i) { return ldg(reinterpret_cast<const uint4 *>(base + (uint32_t)i * 16u)); };
    if (mode == 0) {
        const uint4 c = quad(tp.c);
        decode_point(c.x, tp.ok, unit, o.d[0], o.d[1], o.d[2]);
        decode_point(c.y, tp.ok, unit, o.n[0], o.n[1], o.n[2]);
        decode_point(c.z, tp.ok, unit, o.s[0], o.s[1], o.s[2]);
        decode_point(c.w, tp.ok, unit, o.g, gg, gb);
    } else {
        const uint4 c = quad(tp.c), n0 = quad(tp.n[0]), n1 = quad(tp.n[1]), n2 = quad(tp.n[2]), n3 = quad(tp.n[3]);
        const uint32_t sd[4] = { n0.x, n1.x, n2.x, n3.x }, sn[4] = { n0.y, n1.y, n2.y, n3.y };
        const uint32_t ss[4] = { n0.z, n1.z, n2.z, n3.z }, sg[4] = { n0.w, n1.w, n2.w, n3.w };
        decode_linear(sd, c.x, tp.ok, o.d[0], o.d[1], o.d[2]);
        decode_linear(sn, c.y, tp.ok, o.n[0], o.n[1], o.n[2]);
        decode_linear(ss, c.z, tp.ok, o.s[0], o.s[1], o.s[2]);
        decode_linear(sg, c.w, tp.ok, o.g, gg, gb);
    }
    (void)gg; (void)gb;
    return o;
}

// c / 255.f for a byte value c: the product with the rounded reciprocal and
// one residual correction (fma), which gives the correctly rounded quotient
// for every c in [0, 255].
__device__ __forceinline__ float byte_unit(uint32_t c)
{
    const float rev = 1.0f / 255.f;
    const float f = (float)c;
    const float q = f * rev;
    return __builtin_fmaf(__builtin_fmaf(-q, 255.f, f), rev, q);
}

// The lighting of ShadePixelUnsafe (CCR/CUDAROPs.cu:472-514) on the sampled
// maps c and the interpolated normal n, tangent t (not normalized, as the
// reference leaves it) and view direction v (normalized here), packed to
// RGBA8: normal mapping in the frame {t, -Cross(t, n), n}, Lambert with the
// unclamped observed area (ClampFast(x, 0, x) returns x), the reference's
// Reflect(L, N) = -L + 2 (L.N) N, powf(clamp01(v.R), gloss * 25), MaxToOne.
// Fused multiply-adds and the hardware rsqrt / rcp / exp2 / log2 keep the
// result within 1/255 of the reference formula.
__device__ __forceinline__ uint32_t light_pixel(const MapColours &c, float nx, float ny, float nz, float tx, float ty,
                                                float tz, float vx, float vy, float vz)
{
    const float vl = __builtin_fmaf(vz, vz, __builtin_fmaf(vy, vy, vx * vx));
    const float vinv = vl > 0.f ? __builtin_amdgcn_rsqf(vl) : 0.f;
    vx *= vinv; vy *= vinv; vz *= vinv;
    // binormal = -Cross(tan, n); N = {tan, binormal, n} * (2 * sample - 1)
    const float cx = __builtin_fmaf(ty, nz, -tz * ny), cy = __builtin_fmaf(tz, nx, -tx * nz);
    const float cz = __builtin_fmaf(tx, ny, -ty * nx);
    const float fx = __builtin_fmaf(2.f, c.n[0], -1.f), fy = __builtin_fmaf(2.f, c.n[1], -1.f);
    const float fz = __builtin_fmaf(2.f, c.n[2], -1.f);
    const float Nx = __builtin_fmaf(nx, fz, __builtin_fmaf(-cx, fy, tx * fx));
    const float Ny = __builtin_fmaf(ny, fz, __builtin_fmaf(-cy, fy, ty * fx));
    const float Nz = __builtin_fmaf(nz, fz, __builtin_fmaf(-cz, fy, tz * fx));
    const float Lx = 0.577f, Ly = -0.577f, Lz = -0.577f;
    const float LN = __builtin_fmaf(Lz, Nz, __builtin_fmaf(Ly, Ny, Lx * Nx));
    // observed area -N.L (ClampFast(x, 0, x) leaves x unchanged) * intensity / pi
    const float oa = -LN * (7.0f / 3.14159265358979323846f);
    const float twoDot = 2.f * LN;
    const float rx = __builtin_fmaf(twoDot, Nx, -Lx), ry = __builtin_fmaf(twoDot, Ny, -Ly);
    const float rz = __builtin_fmaf(twoDot, Nz, -Lz);
    float angle = __builtin_fmaf(vz, rz, __builtin_fmaf(vy, ry, vx * rx));
    angle = angle < 0.f ? 0.f : angle;
    angle = angle > 1.f ? 1.f : angle;
    // powf(angle, gloss * 25) via exp2(e * log2(angle)); pow(x, 0) = 1
    const float e = c.g * 25.f;
    angle = e == 0.f ? 1.f : (angle > 0.f ? __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(angle)) : 0.f);
    float outr = __builtin_fmaf(c.s[0], angle, __builtin_fmaf(c.d[0], oa, 0.05f));
    float outg = __builtin_fmaf(c.s[1], angle, __builtin_fmaf(c.d[1], oa, 0.05f));
    float outb = __builtin_fmaf(c.s[2], angle, __builtin_fmaf(c.d[2], oa, 0.05f));
    const float mgb = outg < outb ? outb : outg;
    const float mx = outr < mgb ? mgb : outr;
    if (mx > 1.f) {
        const float rev = __builtin_amdgcn_rcpf(mx);
        outr *= rev; outg *= rev; outb *= rev;
    }
    return pack_rgba(outr, outg, outb);
}

// light_pixel in the reference's own operation order (no fused
// multiply-adds; division and sqrt correctly rounded; powf as in
// light_pixel), for
// triangles whose interpolated attributes can be ill-conditioned (setup marks
// them: a vertex near or behind the camera plane, or large attributes): there
// the fused forms can drift by more than 1/255 from the reference through
// cancellation, while this form gives the CPU path's colour up to the powf
// approximation's 1e-6.  On
// the others every quantity of the lighting stays O(1) and the fused forms
// are within about 1e-5 of it.
__device__ __forceinline__ uint32_t light_pixel_exact(const MapColours &c, float nx, float ny, float nz, float tx,
                                                      float ty, float tz, float vx, float vy, float vz)
{
    normalize3(vx, vy, vz);
    // binormal = -Cross(tan, n)
    const float bx = -(ty * nz - tz * ny), by = -(tz * nx - tx * nz), bz = -(tx * ny - ty * nx);
    const float fx = 2.f * c.n[0] - 1.f, fy = 2.f * c.n[1] - 1.f, fz = 2.f * c.n[2] - 1.f;
    const float Nx = tx * fx + bx * fy + nx * fz;
    const float Ny = ty * fx + by * fy + ny * fz;
    const float Nz = tz * fx + bz * fy + nz * fz;
    const float Lx = 0.577f, Ly = -0.577f, Lz = -0.577f;
    float oa = (-Nx) * Lx + (-Ny) * Ly + (-Nz) * Lz;
    oa /= 3.14159265358979323846f;
    oa *= 7.0f;
    const float twoDot = 2.f * (Lx * Nx + Ly * Ny + Lz * Nz);
    const float rx = -Lx + twoDot * Nx, ry = -Ly + twoDot * Ny, rz = -Lz + twoDot * Nz;
    float angle = vx * rx + vy * ry + vz * rz;
    angle = angle < 0.f ? 0.f : angle;
    angle = angle > 1.f ? 1.f : angle;
    // powf(angle, gloss * 25) as exp2(e * log2(angle)) on the hardware units:
    // for angle in [0, 1] and e <= 25 its error is below 1e-6 absolute
    const float e = c.g * 25.f;
    angle = e == 0.f ? 1.f : (angle > 0.f ? __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(angle)) : 0.f);
    float outr = (0.05f + c.d[0] * oa) + c.s[0] * angle;
    float outg = (0.05f + c.d[1] * oa) + c.s[1] * angle;
    float outb = (0.05f + c.d[2] * oa) + c.s[2] * angle;
    const float mgb = outg < outb ? outb : outg;
    const float mx = outr < mgb ? mgb : outr;
    if (mx > 1.f) {
        const float rev = 1.f / mx;
        outr *= rev; outg *= rev; outb *= rev;
    }
    return pack_rgba(outr, outg, outb);
}

// PS_Kernel + ShadePixelUnsafe (CCR/CUDARenderer.cu:833-897, CCR/CUDAROPs.cu:472-514)
// with the attribute interpolation of RasterizePixelAtomic (CCR/CUDAROPs.cu:347-380)
// for pixel (x, y) of the triangle whose setup record quads are q0, q1, q3.
// The barycentrics, 1/w and the texture coordinates are computed exactly as
// in the reference (so the same texels are fetched).  kExact = false: the
// other attributes and the lighting terms use fused multiply-adds and the
// hardware reciprocal / rsqrt / exp2 / log2, which keeps RGBA8 within 1/255
// of the reference formula; kExact = true (triangles setup marked as
// ill-conditioned) follows the reference's operation order throughout.  A
// pixel's data is a chain of dependent gathers (vis, triangle record, vertex
// attributes and texture table, texels); every level's loads are issued
// together, with no branch between them, and range checks are applied to
// the loaded values.  When all of a wave's pixels belong to one mesh its
// texture table is read with scalar loads, and when its four maps have one
// size the texel indices are computed once and the texels come from the
// maps' interleaved copy (one 16-byte load per tap).
template <bool kLinear, bool kExact, bool kSplit>
__device__ __forceinline__ void shade_body(const DevMesh *__restrict__ meshes, const JobDesc &J, const float *unit,
                                           uint32_t x, uint32_t y, const uint4 q0, const uint4 q1, const uint4 q3)
{
    const FrameParams &fp = J.fp;
    constexpr int mode = kLinear ? 1 : 0;
    const float *__restrict__ depth = J.depth;
    const float4 *__restrict__ attr = J.attr;
    uint32_t *__restrict__ rgba = J.rgba;
    const size_t pid = (size_t)y * fp.width + x;
    const float sx0 = __uint_as_float(q0.x), sy0 = __uint_as_float(q0.y);
    const float sx1 = __uint_as_float(q0.z), sy1 = __uint_as_float(q0.w);
    const float sx2 = __uint_as_float(q1.x), sy2 = __uint_as_float(q1.y);
    const float area = __uint_as_float(q1.w);
    // the three vertices' attributes times 1/w: read as stored, or (views
    // sharing a scene) formed here from the view-independent part and the
    // view's 1/w, the same binary32 products the vertex pass forms otherwise
    float4 a00, a10, a20, a01, a11, a21, a02, a12, a22;
    const float4 *__restrict__ sattr = J.sattr;
    if constexpr (kSplit) {
        const float4 *S0 = sattr + (size_t)q3.x * 2, *S1 = sattr + (size_t)q3.y * 2, *S2 = sattr + (size_t)q3.z * 2;
        const float4 s00 = ldg(S0), s10 = ldg(S1), s20 = ldg(S2);
        const float4 s01 = ldg(S0 + 1), s11 = ldg(S1 + 1), s21 = ldg(S2 + 1);
        const float4 v0 = ldg(attr + q3.x), v1 = ldg(attr + q3.y), v2 = ldg(attr + q3.z);
        auto q0of = [](float4 a, float4 b, float4 vw) {
            return make_float4(a.x * vw.x, a.y * vw.x, vw.x, a.z * vw.x);
        };
        auto q1of = [](float4 a, float4 b, float4 vw) {
            return make_float4(a.w * vw.x, b.x * vw.x, b.y * vw.x, b.z * vw.x);
        };
        auto q2of = [](float4 a, float4 b, float4 vw) { return make_float4(b.w * vw.x, vw.y, vw.z, vw.w); };
        a00 = q0of(s00, s01, v0); a10 = q0of(s10, s11, v1); a20 = q0of(s20, s21, v2);
        a01 = q1of(s00, s01, v0); a11 = q1of(s10, s11, v1); a21 = q1of(s20, s21, v2);
        a02 = q2of(s00, s01, v0); a12 = q2of(s10, s11, v1); a22 = q2of(s20, s21, v2);
    } else {
        const float4 *A0 = attr + (size_t)q3.x * 3, *A1 = attr + (size_t)q3.y * 3, *A2 = attr + (size_t)q3.z * 3;
        a00 = ldg(A0); a10 = ldg(A1); a20 = ldg(A2);
        a01 = ldg(A0 + 1); a11 = ldg(A1 + 1); a21 = ldg(A2 + 1);
        a02 = ldg(A0 + 2); a12 = ldg(A1 + 2); a22 = ldg(A2 + 2);
    }
    const uint32_t me = q3.w & 0x7FFFFFFFu;  // bit 31: the exact-shading mark
    const uint32_t mi = me - J.mesh0 < fp.num_meshes ? me : J.mesh0;  // entry in the launch's mesh table
    const uint32_t mi0 = __builtin_amdgcn_readfirstlane(mi);
    const bool uniform = __ballot(mi != mi0) == 0;
    // the table of a wave-uniform mesh stays in scalar registers (stex); a
    // per-lane one is read only where the maps are sampled, after the
    // attributes have been interpolated
    DevTexture stex[4] = {};
    const uint4 *texq = nullptr;
    int has_diffuse;
    if (uniform) {
        const DevMesh &M = meshes[mi0];
#pragma unroll
        for (int k = 0; k < 4; ++k) stex[k] = M.tex[k];
        texq = M.texq;
        has_diffuse = M.has_diffuse;
    } else {
        has_diffuse = ldg(&meshes[mi].has_diffuse);
    }
    const float px = (float)x, py = (float)y;
    float w2 = (px - sx0) * (sy1 - sy0) - (py - sy0) * (sx1 - sx0);
    float w0 = (px - sx1) * (sy2 - sy1) - (py - sy1) * (sx2 - sx1);
    float w1 = (px - sx2) * (sy0 - sy2) - (py - sy2) * (sx0 - sx2);
    w0 = w0 / area; w1 = w1 / area; w2 = w2 / area;
    const float i0 = a00.z, i1 = a10.z, i2 = a20.z;
    const float wI = 1.f / (i0 * w0 + i1 * w1 + i2 * w2);
    // attributes arrive multiplied by their vertex's 1/w; u and v are formed
    // in the reference's order, the rest with fused multiply-adds unless kExact
#define RAST_INTERP(c0, c1, c2) (w0 * (c0) + w1 * (c1) + w2 * (c2))
#define RAST_FINTERP(c0, c1, c2) \
    (kExact ? RAST_INTERP(c0, c1, c2) : __builtin_fmaf(w2, (c2), __builtin_fmaf(w1, (c1), w0 * (c0))))
    const float u = RAST_INTERP(a00.x, a10.x, a20.x) * wI;
    const float v = RAST_INTERP(a00.y, a10.y, a20.y) * wI;
    const float nx = RAST_FINTERP(a00.w, a10.w, a20.w) * wI;
    const float ny = RAST_FINTERP(a01.x, a11.x, a21.x) * wI;
    const float nz = RAST_FINTERP(a01.y, a11.y, a21.y) * wI;
    if (fp.visualisation == 1) {  // depth colour, Remap(z, 0.985, 1)
        stg(rgba + pid, (uint32_t)(pack_rgba((ldg(depth + pid) - 0.985f) / (1.f - 0.985f), 0.f, 0.f)));
        return;
    }
    if (fp.visualisation == 2) {  // normal colour
        stg(rgba + pid, (uint32_t)(pack_rgba(nx, ny, nz)));
        return;
    }
    if (!has_diffuse) {  // vertex colour (perspective-weighted, not x w)
        // the mesh's input colour (floats 11-13 of the vertex) times the
        // vertex's stored 1/w: the binary32 product the interpolation takes
        const float *iv = ldg(&meshes[mi].ivert);
        const uint32_t vb = ldg(&meshes[mi].vert_base);
        auto colour = [&](uint32_t gv, float iw) {
            const float *p = iv + (size_t)(gv - vb) * 14;
            return make_float4(ldg(p + 11) * iw, ldg(p + 12) * iw, ldg(p + 13) * iw, 0.f);
        };
        const float4 a03 = colour(q3.x, a00.z), a13 = colour(q3.y, a10.z), a23 = colour(q3.z, a20.z);
        stg(rgba + pid, (uint32_t)(pack_rgba(RAST_FINTERP(a03.x, a13.x, a23.x), RAST_FINTERP(a03.y, a13.y, a23.y),
                                             RAST_FINTERP(a03.z, a13.z, a23.z))));
        return;
    }
    // level 4: the texels of all four maps
    MapColours c;
    if (!uniform) {
        DevTexture vtex[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) vtex[k] = ldg(meshes[mi].tex + k);
        c = sample_maps<false>(vtex, mode, u, v, unit);
    } else if (texq) {
        c = sample_quads(texq, stex[0], mode, u, v, unit);
    } else if (stex[1].w == stex[0].w && stex[2].w == stex[0].w && stex[3].w == stex[0].w &&
             stex[1].h == stex[0].h && stex[2].h == stex[0].h && stex[3].h == stex[0].h)
        c = sample_maps<true>(stex, mode, u, v, unit);
    else
        c = sample_maps<false>(stex, mode, u, v, unit);
    const float tx = RAST_FINTERP(a01.z, a11.z, a21.z);
    const float ty = RAST_FINTERP(a01.w, a11.w, a21.w);
    const float tz = RAST_FINTERP(a02.x, a12.x, a22.x);
    const float vx = RAST_FINTERP(a02.y, a12.y, a22.y);
    const float vy = RAST_FINTERP(a02.z, a12.z, a22.z);
    const float vz = RAST_FINTERP(a02.w, a12.w, a22.w);
#undef RAST_INTERP
#undef RAST_FINTERP
    stg(rgba + pid, kExact ? light_pixel_exact(c, nx, ny, nz, tx, ty, tz, vx, vy, vz)
                           : light_pixel(c, nx, ny, nz, tx, ty, tz, vx, vy, vz));
}

// Shades pixel (x, y) showing triangle id (clear colour for none): reads the
// triangle's setup record and runs the fused or the exact shading body as
// setup marked the triangle (two separate code paths, so the rare exact one
// does not raise the register count of the fused one).
template <bool kLinear, bool kSplit>
__device__ __forceinline__ void shade_pixel(const DevMesh *__restrict__ meshes, const JobDesc &J, const float *unit,
                                            uint32_t x, uint32_t y, uint32_t id)
{
    const FrameParams &fp = J.fp;
    if (id >= fp.num_tris) {
        stg(J.rgba + (size_t)y * fp.width + x, (uint32_t)(fp.visualisation == 0 ? fp.clear_rgba : 0xFF000000u));
        return;
    }
    const uint4 *rp = reinterpret_cast<const uint4 *>(J.trirec + (size_t)id * kTriRecWords);
    const uint4 q0 = ldg(rp), q1 = ldg(rp + 1), q3 = ldg(rp + 3);  // q2 holds z2, z1 and the bbox
    if (q3.w >> 31) shade_body<kLinear, true, kSplit>(meshes, J, unit, x, y, q0, q1, q3);
    else shade_body<kLinear, false, kSplit>(meshes, J, unit, x, y, q0, q1, q3);
}


// A block shades 64 x kShadeRows pixels: lane x of wave w takes the pixels
// (x, w), (x, w + 4), ...  The vis reads of all of them go out first, so each
// further pixel of a lane waits for three dependent gathers instead of four.
template <bool kLinear, bool kSplit>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kSplit && !kLinear ? 8 : 1))) void k_shade(const DevMesh *__restrict__ meshes, const JobDesc *__restrict__ jobs,
                                               const uint16_t *__restrict__ table)
{
    __shared__ float unit[256];
    unit[threadIdx.x] = byte_unit(threadIdx.x);
    __syncthreads();
    uint32_t local;
    const JobDesc &J = jobs[job_of(table, jobs, ST_SHADE, local)];
    const FrameParams &fp = J.fp;
    // another instance shades this frame: the other sampling mode's, or the
    // one for the other attribute layout
    if ((fp.sample_state != 0) != kLinear || (J.sattr != nullptr) != kSplit) return;
    // the frame's counters are final once resolve has run: the job's first
    // block copies them to the renderer's pinned host block (k_publish does
    // this for jobs that are not shaded)
    if (local == 0) {
        __shared__ uint32_t red[4];
        publish_counters<4>(J, red);
    }
    constexpr int P = kShadeRows / 4;
    const uint32_t bw = (fp.width + 63u) / 64u;
    const uint32_t x = (local % bw) * 64u + (threadIdx.x & 63u);
    const uint32_t y0 = (local / bw) * (uint32_t)kShadeRows + (threadIdx.x >> 6);
    if (x >= fp.width || y0 >= fp.height) return;
    uint32_t ids[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const uint32_t y = y0 + 4u * k;
        ids[k] = y < fp.height ? ldg(J.vis + (size_t)y * fp.width + x) : 0u;
    }
#pragma unroll
    for (int k = 0; k < P; ++k)
        if (y0 + 4u * k < fp.height) shade_pixel<kLinear, kSplit>(meshes, J, unit, x, y0 + 4u * k, ids[k]);
}

// The persistent form of the resolve stage (the reference's TODO at
// CCR/CUDARenderer.cu:284-285): a grid sized to the workgroups the GPU holds
// at once, each taking the next (job, tile) of the stage from one counter
// over all jobs of the launch until none is left; with SHADE, each tile is
// shaded by the workgroup that resolved it (no visibility re-read, no
// shading launch).  The counter is zeroed by the launch's staging upload.
// Every workgroup leaves once the counter has passed the item count, which
// each fetch increments.
template <int T, int NT, int STAGE, int SHADE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(SHADE ? 3 : 4))) void k_resolve_persistent(
                                                           const DevMesh *__restrict__ meshes,
                                                           const JobDesc *__restrict__ jobs,
                                                           const uint16_t *__restrict__ table,
                                                           uint32_t *__restrict__ work, uint32_t items)
{
    __shared__ uint32_t item_s;
    __shared__ float unit[SHADE ? 256 : 1];
    if constexpr (SHADE != 0) {
        for (int c = (int)threadIdx.x; c < 256; c += NT) unit[c] = byte_unit((uint32_t)c);
    }
    for (;;) {
        // The barrier first: every lane has finished the previous tile (and
        // read item_s) before lane 0 fetches.  It also keeps the compiler
        // from merging the fetch with the lane-0 store that ends the previous
        // tile: merged, the loop is structurized so that the other lanes run
        // ahead into the barrier below while lane 0 waits for its fetch, and
        // they read the old item again, without end.
        __syncthreads();
        if (threadIdx.x == 0) item_s = atomicAdd(work, 1u);
        __syncthreads();
        // the same for every lane, and readfirstlane says so: a scalar exit
        const uint32_t item = __builtin_amdgcn_readfirstlane(item_s);
        if (item >= items) break;
        uint32_t j = 0, local = item;
        if (table) {
            j = __builtin_amdgcn_readfirstlane((uint32_t)table[item]);
            local = item - jobs[j].block0[STAGE];
        }
        resolve_tile<T, NT, STAGE, SHADE>(jobs[j], local, meshes, unit);
    }
}

// DrawTextureKernel / DrawTextureGlobalKernel (CCR/CUDARenderer.cu:1565-1617):
// one texture over the whole frame, one pixel per thread (vis none, depth 1).
__global__ __launch_bounds__(256) void k_draw_texture(uint32_t *__restrict__ vis, float *__restrict__ depth,
                                                      uint32_t *__restrict__ rgba, uint32_t width, uint32_t height,
                                                      const uint32_t *__restrict__ texels, uint32_t tw, uint32_t th,
                                                      int mode, int stretch, int sample_state)
{
    const uint32_t x = blockIdx.x * 64u + (threadIdx.x & 63u);
    const uint32_t y = blockIdx.y * 4u + (threadIdx.x >> 6);
    if (x >= width || y >= height) return;
    const size_t pid = (size_t)y * width + x;
    stg(vis + pid, kVisNone);
    stg(depth + pid, 1.0f);
    uint32_t out;
    if (mode == 0) {
        // normalized coordinates, point filter, wrap addressing
        const float u = (float)x / (float)width, v = (float)y / (float)height;
        int ix = (int)floorf(u * (float)tw), iy = (int)floorf(v * (float)th);
        ix %= (int)tw;
        iy %= (int)th;
        ix += ix < 0 ? (int)tw : 0;
        iy += iy < 0 ? (int)th : 0;
        out = ldg(texels + (size_t)iy * tw + (uint32_t)ix);
    } else {
        float u = (float)x, v = (float)y;
        if (stretch) {
            u /= (float)width;
            v /= (float)height;
        } else {
            u /= (float)tw;
            v /= (float)th;
        }
        DevTexture t;
        t.texels = texels;
        t.w = tw;
        t.h = th;
        const Taps tp = texel_taps(t, u, v);
        float r, g, b;
        if (sample_state == 0) {
            const uint32_t c = ldg(texels + tp.c);
            r = tp.ok ? byte_unit(c & 0xFFu) : 1.f;
            g = tp.ok ? byte_unit((c >> 8) & 0xFFu) : 0.f;
            b = tp.ok ? byte_unit((c >> 16) & 0xFFu) : 1.f;
        } else {
            uint32_t n[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) n[i] = ldg(texels + tp.n[i]);
            decode_linear(n, ldg(texels + tp.c), tp.ok, r, g, b);
        }
        out = pack_rgba(r, g, b);
    }
    stg(rgba + pid, out);
}

// Copies the counters of every job that is not shaded to its renderer's
// pinned host block (one workgroup per job), replacing a device-to-host copy
// per job; the shading kernel publishes the others.
__global__ __launch_bounds__(256) void k_publish(const JobDesc *__restrict__ jobs, int all)
{
    __shared__ uint32_t red[4];
    const JobDesc &J = jobs[blockIdx.x];
    if (J.nblocks[ST_SHADE] && !all) return;
    publish_counters<4>(J, red);
}

inline uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

hipError_t launch_interleave_maps(const uint32_t *d, const uint32_t *n, const uint32_t *sp, const uint32_t *g,
                                  uint4 *out, size_t count, hipStream_t s)
{
    const size_t blocks = (count + 255) / 256;
    hipLaunchKernelGGL(k_interleave_maps, dim3(blocks < 4096 ? (uint32_t)blocks : 4096u), dim3(256), 0, s, d, n, sp, g,
                       out, count);
    return hipGetLastError();
}

hipError_t launch_clear_keys(unsigned long long *keys, size_t n, hipStream_t s)
{
    const size_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_clear_keys, dim3(blocks < 2048 ? (uint32_t)blocks : 2048u), dim3(256), 0, s, keys, n);
    return hipGetLastError();
}

hipError_t launch_vertex(const Launch &L, int with_capacity, hipStream_t s)
{
    const uint32_t cap_blocks = L.grid[ST_CAP];
    const uint32_t vs_blocks = with_capacity == 2 ? 0u : L.vs_blocks;  // 2: the capacity stage alone
    const uint32_t init_blocks = with_capacity == 2 ? 0u : L.grid[ST_INIT];
    const uint16_t *init_table = L.table[ST_INIT];
    if (!with_capacity || !cap_blocks) {
        if (vs_blocks + init_blocks == 0) return hipSuccess;
        hipLaunchKernelGGL(k_vertex<0>, dim3(init_blocks + vs_blocks), dim3(256), 0, s, L.meshes, L.vs_block_group,
                           L.vs_groups, L.vs_entries, L.jobs, (const uint16_t *)nullptr, 0u, init_table, init_blocks);
        return hipGetLastError();
    }
    const dim3 g(cap_blocks + init_blocks + vs_blocks), b(256);
    // one instance per hidden width (a launch's learned jobs share one);
    // heuristic-only launches take the smallest
    switch (L.model_hidden) {
    case 32:
        hipLaunchKernelGGL(k_vertex<32>, g, b, 0, s, L.meshes, L.vs_block_group, L.vs_groups, L.vs_entries, L.jobs,
                           L.table[ST_CAP], cap_blocks, init_table, init_blocks);
        break;
    case 64:
        hipLaunchKernelGGL(k_vertex<64>, g, b, 0, s, L.meshes, L.vs_block_group, L.vs_groups, L.vs_entries, L.jobs,
                           L.table[ST_CAP], cap_blocks, init_table, init_blocks);
        break;
    default:
        hipLaunchKernelGGL(k_vertex<16>, g, b, 0, s, L.meshes, L.vs_block_group, L.vs_groups, L.vs_entries, L.jobs,
                           L.table[ST_CAP], cap_blocks, init_table, init_blocks);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_setup(const Launch &L, hipStream_t s)
{
    if (L.tri_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_setup, dim3(L.tri_blocks), dim3(256), 0, s, L.meshes, L.tri_block_mesh, L.jobs, L.tri_blocks);
    return hipGetLastError();
}

hipError_t launch_scan(const Launch &L, hipStream_t s)
{
    if (L.exact_jobs == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scan, dim3(L.njobs), dim3(1024), 0, s, L.jobs);
    return hipGetLastError();
}

hipError_t launch_fill(const Launch &L, hipStream_t s)
{
    if (!L.grid[ST_FILL]) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(L.grid[ST_FILL]), dim3(256), 0, s, L.jobs, L.table[ST_FILL]);
    return hipGetLastError();
}

hipError_t launch_fallback(const Launch &L, hipStream_t s)
{
    if (!L.grid[ST_FBALLOC]) return hipSuccess;
    hipLaunchKernelGGL(k_fallback_alloc, dim3(L.grid[ST_FBALLOC]), dim3(256), 0, s, L.jobs, L.table[ST_FBALLOC]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fallback_fill, dim3(L.grid[ST_FBFILL]), dim3(256), 0, s, L.jobs, L.table[ST_FBFILL]);
    return hipGetLastError();
}

template <int T, int NT, int STAGE>
static hipError_t launch_resolve_persistent(const Launch &L, hipStream_t s)
{
    const uint32_t items = L.grid[STAGE];
    const uint32_t g = items < L.resident[STAGE] ? items : L.resident[STAGE];
    uint32_t *work = L.work + STAGE;
    switch (L.fused_shade) {
    case 1:
        hipLaunchKernelGGL((k_resolve_persistent<T, NT, STAGE, 1>), dim3(g), dim3(NT), 0, s, L.meshes, L.jobs,
                           L.table[STAGE], work, items);
        break;
    case 2:
        hipLaunchKernelGGL((k_resolve_persistent<T, NT, STAGE, 2>), dim3(g), dim3(NT), 0, s, L.meshes, L.jobs,
                           L.table[STAGE], work, items);
        break;
    default:
        hipLaunchKernelGGL((k_resolve_persistent<T, NT, STAGE, 0>), dim3(g), dim3(NT), 0, s, L.meshes, L.jobs,
                           L.table[STAGE], work, items);
        break;
    }
    return hipGetLastError();
}

int resolve_resident_blocks(int stage, int fused_shade)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    hipError_t e = hipErrorInvalidValue;
#define RAST_OCC(T, NT, ST)                                                                                      \
    switch (fused_shade) {                                                                                     \
    case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_resolve_persistent<T, NT, ST, 1>, NT, 0); break; \
    case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_resolve_persistent<T, NT, ST, 2>, NT, 0); break; \
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_resolve_persistent<T, NT, ST, 0>, NT, 0); break; \
    }
    switch (stage) {
    case ST_RES8: RAST_OCC(8, 64, ST_RES8); break;
    case ST_RES16: RAST_OCC(16, 256, ST_RES16); break;
    case ST_RES32: RAST_OCC(32, 256, ST_RES32); break;
    default: return 0;
    }
#undef RAST_OCC
    if (e != hipSuccess || per_cu <= 0) return 0;
    return per_cu * cus;
}

hipError_t launch_resolve(const Launch &L, int stage, hipStream_t s)
{
    const uint32_t g = L.grid[stage];
    if (!g) return hipSuccess;
    if (L.resolve_mode != 0 && L.resident[stage]) {
        switch (stage) {
        case ST_RES8: return launch_resolve_persistent<8, 64, ST_RES8>(L, s);
        case ST_RES16: return launch_resolve_persistent<16, 256, ST_RES16>(L, s);
        case ST_RES32: return launch_resolve_persistent<32, 256, ST_RES32>(L, s);
        default: return hipErrorInvalidValue;
        }
    }
    switch (stage) {
    case ST_RES8: hipLaunchKernelGGL((k_resolve<8, 64, ST_RES8>), dim3(g), dim3(64), 0, s, L.jobs, L.table[ST_RES8]); break;
    case ST_RES16:
        hipLaunchKernelGGL((k_resolve<16, 256, ST_RES16>), dim3(g), dim3(256), 0, s, L.jobs, L.table[ST_RES16]);
        break;
    case ST_RES32:
        hipLaunchKernelGGL((k_resolve<32, 256, ST_RES32>), dim3(g), dim3(256), 0, s, L.jobs, L.table[ST_RES32]);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The launch's staging block, read by the GPU straight from its pinned
// (coherent) host copy: one 16-byte word per lane.  A kernel of its own
// instead of a host-to-device copy on the stream, which the runtime does as
// a blit kernel behind extra synchronization (4K frame 1.3 us shorter).
__global__ __launch_bounds__(256) void k_upload(uint4 *__restrict__ dst, const uint4 *__restrict__ src, uint32_t n)
{
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) stg(dst + i, ldg(src + i));
}

hipError_t launch_upload(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    if (bytes % 16 != 0 || bytes / 16 > 0xFFFFFFFFu) return hipErrorInvalidValue;
    const uint32_t n = (uint32_t)(bytes / 16);
    uint32_t g = (n + 255u) / 256u;
    g = g < 1u ? 1u : (g > 64u ? 64u : g);
    hipLaunchKernelGGL(k_upload, dim3(g), dim3(256), 0, s, (uint4 *)dst, (const uint4 *)src, n);
    return hipGetLastError();
}

hipError_t launch_shade(const Launch &L, hipStream_t s)
{
    if (!L.grid[ST_SHADE] || L.fused_shade) return hipSuccess;
    // one instance per sampling mode and attribute layout present: a mode's
    // code alone keeps the register count (and so the waves per SIMD) of that
    // instance
    if (L.shade_modes & 1u)
        hipLaunchKernelGGL((k_shade<false, false>), dim3(L.grid[ST_SHADE]), dim3(256), 0, s, L.meshes, L.jobs,
                           L.table[ST_SHADE]);
    if (L.shade_modes & 2u)
        hipLaunchKernelGGL((k_shade<true, false>), dim3(L.grid[ST_SHADE]), dim3(256), 0, s, L.meshes, L.jobs,
                           L.table[ST_SHADE]);
    if (L.shade_modes & 4u)
        hipLaunchKernelGGL((k_shade<false, true>), dim3(L.grid[ST_SHADE]), dim3(256), 0, s, L.meshes, L.jobs,
                           L.table[ST_SHADE]);
    if (L.shade_modes & 8u)
        hipLaunchKernelGGL((k_shade<true, true>), dim3(L.grid[ST_SHADE]), dim3(256), 0, s, L.meshes, L.jobs,
                           L.table[ST_SHADE]);
    return hipGetLastError();
}

hipError_t launch_draw_texture(uint32_t *vis, float *depth, uint32_t *rgba, uint32_t width, uint32_t height,
                               const uint32_t *texels, uint32_t tw, uint32_t th, int mode, bool stretch,
                               int sample_state, hipStream_t s)
{
    hipLaunchKernelGGL(k_draw_texture, dim3(ceil_div(width, 64), ceil_div(height, 4)), dim3(256), 0, s, vis, depth,
                       rgba, width, height, texels, tw, th, mode, stretch ? 1 : 0, sample_state);
    return hipGetLastError();
}

hipError_t launch_publish(const Launch &L, hipStream_t s)
{
    // shading in the resolve pass: every job's counters go out here
    const int all = L.fused_shade != 0;
    if (L.unshaded_jobs == 0 && !all) return hipSuccess;  // every job's shading kernel publishes
    hipLaunchKernelGGL(k_publish, dim3(L.njobs), dim3(256), 0, s, L.jobs, all);
    return hipGetLastError();
}

}  // namespace rast
