// rast_capacity.h -- single-pass queue capacities of a frame's tiles and the
// placement of their queue regions (device code shared by the vertex launch,
// which runs it in workgroups of its own beside the vertex transform, and by
// the capacity model's training-set recorder).
//
// A tile's capacity is a hint: ceil(expm1(pred) * margin) + slack from the
// learned model, or ceil(prev * margin) + slack from the previous frame's
// count (the heuristic), at most the frame's triangle count.  Entries beyond
// a capacity are re-binned exactly, so a capacity never changes the frame.
// The reference has no counterpart: it sizes every queue to a compile-time
// constant (CCR/DEFINES.h:21-23).
#pragma once

#include <hip/hip_runtime.h>

#include "rast_internal.h"
#include "rast_launch.h"

namespace rast {

constexpr int kCapFeatures = kNumFeatures;
static_assert(kCapFeatures == 8, "the layer-1 operand gather assumes two 4-feature steps");
static_assert(kCapacityTilesPerBlock == 256, "one tile per lane of a 256-lane workgroup");

// A tile's single-pass queue capacity from a predicted length c (>= 0, or
// NaN): c + slack, at most `limit` (the frame's triangle count: a tile's
// queue never holds a triangle twice).  Formed in 64 bits, so neither the
// float conversion nor the slack addition can wrap.
__device__ __forceinline__ uint32_t clamp_capacity(float c, uint32_t slack, uint32_t limit)
{
    const uint64_t base = c > 0.f ? (c < (float)limit ? (uint64_t)c : (uint64_t)limit) : 0ull;
    const uint64_t v = base + slack;
    return v < limit ? (uint32_t)v : limit;
}

// Per-tile features (inputs of the capacity model), shared by inference and
// training-set recording, all from the previous frame (the model runs before
// this frame's vertex pass is done): the tile's exact queue length, the mean
// and max of its 3x3 and of its 5x5 neighbourhood (the wider ring follows
// camera motion), its min depth, and the tile position.
__device__ __forceinline__ void tile_features(const FrameParams &fp, uint32_t t, const uint32_t *__restrict__ prev,
                                              const float *__restrict__ prev_depth, float *o)
{
    const int tx = (int)(t % fp.tiles_x), ty = (int)(t / fp.tiles_x);
    // all loads first (tiles outside the grid count as absent), so they are
    // in flight together; then the sums in neighbourhood order
    uint32_t pc[25];
    bool ok[25];
#pragma unroll
    for (int k = 0; k < 25; ++k) {
        const int x = tx + k % 5 - 2, y = ty + k / 5 - 2;
        ok[k] = x >= 0 && y >= 0 && x < (int)fp.tiles_x && y < (int)fp.tiles_y;
        const uint32_t q = ok[k] ? (uint32_t)y * fp.tiles_x + (uint32_t)x : t;
        pc[k] = prev[q];
    }
    float s3 = 0.f, m3 = 0.f, s5 = 0.f, m5 = 0.f;
    int n3 = 0, n5 = 0;
#pragma unroll
    for (int k = 0; k < 25; ++k) {
        if (!ok[k]) continue;
        const float c = (float)pc[k];
        const bool inner = k % 5 >= 1 && k % 5 <= 3 && k / 5 >= 1 && k / 5 <= 3;
        if (inner) {
            s3 += c;
            m3 = c > m3 ? c : m3;
            ++n3;
        }
        s5 += c;
        m5 = c > m5 ? c : m5;
        ++n5;
    }
    o[0] = log1pf((float)pc[12]);
    o[1] = log1pf(s3 / (float)n3);
    o[2] = log1pf(m3);
    o[3] = log1pf(s5 / (float)n5);
    o[4] = log1pf(m5);
    o[5] = prev_depth[t];
    o[6] = ((float)ty + 0.5f) / (float)fp.tiles_y;
    o[7] = ((float)tx + 0.5f) / (float)fp.tiles_x;
}

// LDS floats the capacity workgroup of a model of hidden width H stages:
// W2 and W1 in operand order, then b1, b2 and w3.
template <int H>
constexpr int capacity_lds_floats()
{
    return H ? H * H + H * kCapFeatures + 3 * H : 0;
}

typedef float cap_v4f __attribute__((ext_vector_type(4)));

// The learned capacity of this lane's tile (a workgroup barrier inside: the
// whole workgroup calls it).  The MLP F -> H -> H -> 1 (ReLU) runs on the
// matrix cores as v_mfma_f32_16x16x4f32 (A 16x4: lane l holds row l % 16,
// k = l / 16; B 4x16: lane l holds k = l / 16, column l % 16; D 16x16: lane l
// holds column l % 16, rows 4 (l / 16) + r in register r), each layer
// computed transposed, units x tiles: D1 = W1 X^T, D2 = W2 H1^T.  Lane l of
// D1's 16-unit block jc then holds tile l % 16 and units jc*16 + 4 (l / 16) + r
// in register r, which is the B operand of layer 2 as it stands when layer 2
// sums over the input units in the steps (jc, r), each step taking the four
// units jc*16 + 4q + r (q = 0..3): no activation passes through memory.  The
// weights are staged in LDS in that operand order, once per workgroup.  A
// wave serves 64 tiles (one per lane, whose features the lane computes) in
// four passes of 16; a pass gathers its tiles' features with lane shuffles.
template <int H>
__device__ float learned_prediction(const JobDesc &J, uint32_t t, bool live, float *lds)
{
    static_assert(H % 16 == 0 && H >= 16 && H <= 64, "hidden width 16, 32 or 64");
    constexpr int F = kCapFeatures, NJ = H / 16;
    float *w2p = lds;            // [jo][jc][lane][r] = W2[jo*16 + lane%16][jc*16 + 4 (lane/16) + r]
    float *w1p = w2p + H * H;    // [jc][ks][lane]    = W1[jc*16 + lane%16][4 ks + lane/16]
    float *b1s = w1p + H * F, *b2s = b1s + H, *w3s = b2s + H;
    // parameter layout: W1 [H][F], b1 [H], W2 [H][H], b2 [H], W3 [H], b3
    const float *__restrict__ p = J.model;
    const float *gW1 = p, *gb1 = gW1 + H * F, *gW2 = gb1 + H, *gb2 = gW2 + H * H, *gW3 = gb2 + H;
    const float b3 = p[H * F + H + H * H + H + H];
    // W1 and W2 start at multiples of 16 bytes (offsets 0 and 9 H floats of
    // a 256-byte aligned buffer): 16-byte loads
    for (int e4 = (int)threadIdx.x; e4 < H * H / 4; e4 += 256) {
        const float4 v = reinterpret_cast<const float4 *>(gW2)[e4];
        const int o = 4 * e4 / H, i = 4 * e4 % H;  // i % 4 == 0: the four components are r = 0..3
        reinterpret_cast<float4 *>(w2p)[((o >> 4) * NJ + (i >> 4)) * 64 + (o & 15) + 16 * ((i & 15) >> 2)] = v;
    }
    for (int e4 = (int)threadIdx.x; e4 < H * F / 4; e4 += 256) {
        const float4 v = reinterpret_cast<const float4 *>(gW1)[e4];
        const int u = 4 * e4 / F, f = 4 * e4 % F;  // f % 4 == 0: components are lanes + 0, 16, 32, 48
        float *d = w1p + ((u >> 4) * 2 + f / 4) * 64 + (u & 15);
        d[0] = v.x;
        d[16] = v.y;
        d[32] = v.z;
        d[48] = v.w;
    }
    for (int i = (int)threadIdx.x; i < H; i += 256) {
        b1s[i] = gb1[i];
        b2s[i] = gb2[i];
        w3s[i] = gW3[i];
    }
    float x[F];
    if (live) {
        tile_features(J.fp, t, J.prev_count, J.prev_depth, x);
    } else {
#pragma unroll
        for (int f = 0; f < F; ++f) x[f] = 0.f;
    }
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63u), m = lane & 15, q = lane >> 4;
    const float4 *w2v = reinterpret_cast<const float4 *>(w2p);
    const float4 *b1v = reinterpret_cast<const float4 *>(b1s);
    const float4 *b2v = reinterpret_cast<const float4 *>(b2s);
    const float4 *w3v = reinterpret_cast<const float4 *>(w3s);
    float own = 0.f;
#pragma unroll 1
    for (int pass = 0; pass < 4; ++pass) {
        // the weight reads below do not depend on the pass; hoisted out of
        // the loop they would hold 4 H + 72 registers for the whole launch
        // (the vertex workgroups' occupancy with them), so their indices are
        // made opaque to the compiler in each pass
        int wl = lane, wq = q;
        asm volatile("" : "+v"(wl), "+v"(wq));
        // B of layer 1: feature 4 ks + q of the pass's tile m (lane 16 pass + m)
        float xb[2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int src = 16 * pass + m;
            const float g0 = __shfl(x[4 * ks + 0], src), g1 = __shfl(x[4 * ks + 1], src);
            const float g2 = __shfl(x[4 * ks + 2], src), g3 = __shfl(x[4 * ks + 3], src);
            xb[ks] = q == 0 ? g0 : q == 1 ? g1 : q == 2 ? g2 : g3;
        }
        cap_v4f h1[NJ];
#pragma unroll
        for (int jc = 0; jc < NJ; ++jc) {
            cap_v4f acc = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(w1p[(jc * 2 + ks) * 64 + wl], xb[ks], acc, 0, 0, 0);
            const float4 b = b1v[jc * 4 + wq];  // units jc*16 + 4q + r
            const float bb[4] = { b.x, b.y, b.z, b.w };
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float hv = acc[r] + bb[r];
                h1[jc][r] = hv > 0.f ? hv : 0.f;
            }
        }
        // layer 2 one 16-unit output block at a time (its accumulator alone
        // is live), layer 3 over this lane's four units of the block
        float s = 0.f;
#pragma unroll
        for (int jo = 0; jo < NJ; ++jo) {
            cap_v4f acc2 = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
            for (int jc = 0; jc < NJ; ++jc) {
                const float4 a = w2v[(jo * NJ + jc) * 64 + wl];
                acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, h1[jc][0], acc2, 0, 0, 0);
                acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, h1[jc][1], acc2, 0, 0, 0);
                acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, h1[jc][2], acc2, 0, 0, 0);
                acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, h1[jc][3], acc2, 0, 0, 0);
            }
            const float4 b = b2v[jo * 4 + wq], w3 = w3v[jo * 4 + wq];
            const float bb[4] = { b.x, b.y, b.z, b.w }, ww[4] = { w3.x, w3.y, w3.z, w3.w };
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float hv = acc2[r] + bb[r];
                s += ww[r] * (hv > 0.f ? hv : 0.f);
            }
        }
        // the tile's other units sit in the other three lane groups
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (q == pass) own = b3 + s;  // lane 16 pass + m serves tile 16 pass + m of the wave
    }
    return own;
}

// Queue regions of a wave's 64 tiles: one atomic on the frame's 64-bit
// allocation top per wave reserves the sum of the capacities, and each tile
// takes its place in that block by the wave's prefix sum.  A region that
// would end past the queue storage keeps only the room that is left (the
// rest of its entries go to the exact fallback); the top keeps the full sum,
// from which the host grows the storage for the next frame.  Regions are
// ordered by wave, not by tile: a tile's queue is read through its own
// offset only.
__device__ __forceinline__ void place_queue_regions(const JobDesc &J, uint32_t t, bool live, uint32_t cap)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t c = live ? cap : 0u;
    uint64_t incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(incl, o);
        if (lane >= (uint32_t)o) incl += y;
    }
    const uint64_t total = __shfl(incl, 63);
    uint64_t base = 0;
    if (total) {  // wave-uniform
        if (lane == 63) base = atomicAdd(&J.counters->capacity_top, (unsigned long long)total);
        base = __shfl(base, 63);
    }
    if (live) {
        const uint64_t start = base + incl - c, limit = J.queue_cap;
        J.tile_off[t] = (uint32_t)(start < limit ? start : limit);
        J.tile_cap[t] = start >= limit ? 0u : (uint32_t)(c < limit - start ? c : limit - start);
    }
}

// One capacity workgroup (256 tiles of one job; block b of the ST_CAP grid).
template <int H>
__device__ __forceinline__ void capacity_block(const JobDesc *__restrict__ jobs, const uint16_t *__restrict__ table,
                                               float *lds)
{
    uint32_t local;
    const JobDesc &J = jobs[job_of(table, jobs, ST_CAP, local)];
    const FrameParams &fp = J.fp;
    const uint32_t n = fp.tiles_x * fp.tiles_y;
    const uint32_t t = local * (uint32_t)kCapacityTilesPerBlock + threadIdx.x;
    const bool live = t < n;
    uint32_t cap = 0;
    if (J.learned) {  // the same for the whole workgroup
        const float out = learned_prediction<H>(J, t, live, lds);
        if (live) cap = clamp_capacity(ceilf(expm1f(out < 20.f ? out : 20.f) * J.margin), J.slack, fp.num_tris);
    } else if (live) {
        cap = clamp_capacity(ceilf((float)J.prev_count[t] * J.margin), J.slack, fp.num_tris);
    }
    place_queue_regions(J, t, live, cap);
}

}  // namespace rast
