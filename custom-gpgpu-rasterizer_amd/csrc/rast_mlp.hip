// rast_mlp.hip -- learned bin-capacity model for gfx950.
//
// A per-tile MLP in -> H -> H -> 1 (ReLU, fp32 weights and accumulation)
// predicts log(1 + queue length) of each screen tile from features available
// before binning: the previous frame's exact queue lengths (the tile and its
// 3x3 and 5x5 neighbourhoods), the previous frame's min depth and the tile
// position: all known before this frame's vertex pass, in whose launch the
// frame-time inference runs (rast_capacity.h).  Capacities
// ceil(expm1(pred) * margin) + slack let binning run in one pass; entries
// beyond a tile's capacity are re-binned exactly (rast_kernels.hip), so the
// prediction never changes the frame.  The reference has no counterpart: it
// sizes every queue to a compile-time constant (CCR/DEFINES.h:21-23).
//
// Training runs on the GPU with the kernels below: forward + backward per
// sample (lane = sample), weight gradients as fixed-order reductions over the
// batch (deterministic), and Adam.  Loss: pinball (quantile tau) on
// target - pred, averaged over the batch; its gradient is checked against
// the CPU oracle.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "rast_capacity.h"
#include "rast_launch.h"
#include "rast_model.h"
#include "rast_state.h"

using namespace rast;

namespace {

constexpr int F = kNumFeatures;

__host__ __device__ inline size_t num_params(int H) { return (size_t)H * F + H + (size_t)H * H + H + H + 1; }

// ---------------------------------------------------------------- features

// Training-set features of a rendered frame (tile_features: rast_capacity.h).
__global__ __launch_bounds__(256) void k_features(FrameParams fp, const uint32_t *__restrict__ prev,
                                                  const float *__restrict__ prev_depth, float *__restrict__ out)
{
    const uint32_t n = fp.tiles_x * fp.tiles_y;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    float f[F];
    tile_features(fp, t, prev, prev_depth, f);
    for (int q = 0; q < F; ++q) out[(size_t)t * F + q] = f[q];
}

__global__ __launch_bounds__(256) void k_targets(uint32_t n, const uint32_t *__restrict__ count,
                                                 float *__restrict__ y)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t < n) y[t] = log1pf((float)count[t]);
}

// ----------------------------------------------------------------- forward

// One sample per lane; weights staged in LDS (broadcast reads).
template <int H>
__device__ __forceinline__ float mlp_forward_lane(const float *__restrict__ w, const float *x, float *h1)
{
    const float *W1 = w, *b1 = W1 + H * F, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + H;
#pragma unroll
    for (int k = 0; k < H; ++k) {
        float a = b1[k];
#pragma unroll
        for (int f = 0; f < F; ++f) a += W1[k * F + f] * x[f];
        h1[k] = a > 0.f ? a : 0.f;
    }
    float out = b3[0];
    for (int j = 0; j < H; ++j) {
        float a = b2[j];
#pragma unroll
        for (int k = 0; k < H; ++k) a += W2[j * H + k] * h1[k];
        out += W3[j] * (a > 0.f ? a : 0.f);
    }
    return out;
}

template <int H>
__global__ __launch_bounds__(256) void k_mlp_infer(uint32_t n, const float *__restrict__ x, const float *__restrict__ p,
                                                   float margin, uint32_t slack, uint32_t *__restrict__ cap,
                                                   float *__restrict__ pred)
{
    constexpr int NP = H * F + H + H * H + H + H + 1;
    __shared__ float w[NP];
    for (int i = threadIdx.x; i < NP; i += 256) w[i] = p[i];
    __syncthreads();
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    float xv[F], h1[H];
#pragma unroll
    for (int f = 0; f < F; ++f) xv[f] = x[(size_t)t * F + f];
    const float o = mlp_forward_lane<H>(w, xv, h1);
    if (pred) pred[t] = o;
    if (cap) cap[t] = clamp_capacity(ceilf(expm1f(o < 20.f ? o : 20.f) * margin), slack, 0xFFFFFFFFu);
}

// ----------------------------------------------------------------- training

// One training step is split over blocks of 64 samples (split-K): wave 0
// runs forward + backward with one sample per lane (weights in LDS) and keeps
// activations / deltas in LDS; then all 256 threads form the block's partial
// weight gradients, which k_reduce_partials sums in a fixed order, so a step
// is deterministic.
template <int H>
__global__ __launch_bounds__(256) void k_train_partial(uint32_t B, uint32_t N, const float *__restrict__ X,
                                                       const float *__restrict__ Y, const uint32_t *__restrict__ idx,
                                                       const float *__restrict__ p, float tau,
                                                       float *__restrict__ partial)
{
    constexpr int NP = H * F + H + H * H + H + H + 1;
    __shared__ float w[NP];
    __shared__ float sx[64][F + 1];
    __shared__ float sh1[64][H + 1], sh2[64][H + 1], sd1[64][H + 1], sd2[64][H + 1];
    __shared__ float sdout[64], sloss[64];
    for (int i = threadIdx.x; i < NP; i += 256) w[i] = p[i];
    __syncthreads();
    const float *W1 = w, *b1 = W1 + H * F, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + H;
    const uint32_t s0 = blockIdx.x * 64u;
    if (threadIdx.x < 64) {
        const uint32_t lane = threadIdx.x, s = s0 + lane;
        float xv[F], h1[H], h2[H], d2[H];
        float g = 0.f, l = 0.f;
        if (s < B) {
            uint32_t src = idx ? idx[s] : s;
            src = src < N ? src : N - 1;
#pragma unroll
            for (int f = 0; f < F; ++f) xv[f] = X[(size_t)src * F + f];
#pragma unroll
            for (int k = 0; k < H; ++k) {
                float a = b1[k];
#pragma unroll
                for (int f = 0; f < F; ++f) a += W1[k * F + f] * xv[f];
                h1[k] = a > 0.f ? a : 0.f;
            }
            float out = b3[0];
            for (int j = 0; j < H; ++j) {
                float a = b2[j];
#pragma unroll
                for (int k = 0; k < H; ++k) a += W2[j * H + k] * h1[k];
                h2[j] = a > 0.f ? a : 0.f;
                out += W3[j] * h2[j];
            }
            const float r = Y[src] - out;
            l = r > 0.f ? tau * r : (tau - 1.f) * r;
            g = (r > 0.f ? -tau : 1.f - tau) / (float)B;
        } else {
#pragma unroll
            for (int f = 0; f < F; ++f) xv[f] = 0.f;
#pragma unroll
            for (int k = 0; k < H; ++k) { h1[k] = 0.f; h2[k] = 0.f; }
        }
#pragma unroll
        for (int j = 0; j < H; ++j) d2[j] = h2[j] > 0.f ? g * W3[j] : 0.f;
        for (int k = 0; k < H; ++k) {
            float a = 0.f;
#pragma unroll
            for (int j = 0; j < H; ++j) a += W2[j * H + k] * d2[j];
            sd1[lane][k] = h1[k] > 0.f ? a : 0.f;
            sh1[lane][k] = h1[k];
        }
#pragma unroll
        for (int j = 0; j < H; ++j) { sd2[lane][j] = d2[j]; sh2[lane][j] = h2[j]; }
#pragma unroll
        for (int f = 0; f < F; ++f) sx[lane][f] = xv[f];
        sdout[lane] = g;
        sloss[lane] = l;
    }
    __syncthreads();
    float *out = partial + (size_t)blockIdx.x * (NP + 1);
    float *gW1 = out, *gb1 = gW1 + H * F, *gW2 = gb1 + H, *gb2 = gW2 + H * H, *gW3 = gb2 + H, *gb3 = gW3 + H;
    for (int o = threadIdx.x; o < H * H; o += 256) {
        const int j = o / H, k = o % H;
        float a = 0.f;
        for (int s = 0; s < 64; ++s) a += sd2[s][j] * sh1[s][k];
        gW2[o] = a;
    }
    for (int o = threadIdx.x; o < H * F; o += 256) {
        const int j = o / F, f = o % F;
        float a = 0.f;
        for (int s = 0; s < 64; ++s) a += sd1[s][j] * sx[s][f];
        gW1[o] = a;
    }
    if (threadIdx.x < H) {
        const int j = threadIdx.x;
        float a1 = 0.f, a2 = 0.f, a3 = 0.f;
        for (int s = 0; s < 64; ++s) {
            a1 += sd1[s][j];
            a2 += sd2[s][j];
            a3 += sdout[s] * sh2[s][j];
        }
        gb1[j] = a1;
        gb2[j] = a2;
        gW3[j] = a3;
    }
    if (threadIdx.x == 255) {
        float a = 0.f, lsum = 0.f;
        for (int s = 0; s < 64; ++s) { a += sdout[s]; lsum += sloss[s]; }
        gb3[0] = a;
        out[NP] = lsum;  // loss partial
    }
}

// grad[i] = sum over blocks of partial[b][i] in block order; loss = mean.
__global__ __launch_bounds__(256) void k_reduce_partials(uint32_t nblocks, uint32_t np, uint32_t B,
                                                         const float *__restrict__ partial, float *__restrict__ grad,
                                                         float *__restrict__ loss)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i > np) return;
    float a = 0.f;
    for (uint32_t b = 0; b < nblocks; ++b) a += partial[(size_t)b * (np + 1) + i];
    if (i < np) grad[i] = a;
    else if (loss) loss[0] = a / (float)B;
}

__global__ __launch_bounds__(256) void k_adam(uint32_t n, float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, float lr, float b1c,
                                              float b2c)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / b1c) / (sqrtf(vi / b2c) + 1e-8f);
}

__global__ __launch_bounds__(256) void k_batch_index(uint32_t B, uint32_t N, uint32_t seed, uint32_t *__restrict__ idx)
{
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= B) return;
    uint32_t x = seed ^ (s * 0x9E3779B9u);
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    idx[s] = (uint32_t)(((uint64_t)x * N) >> 32);
}

uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

// ------------------------------------------------------------------ object

struct rast_binmodel {
    int device = 0;
    int H = 64;
    hipStream_t stream = nullptr;
    DBuf params, grad, adam_m, adam_v;
    DBuf X, Y;            // recorded training set
    uint32_t n = 0, cap = 0;
    DBuf idx, partial, loss, io_x, io_y, io_p;  // partial: per 64-sample block gradients + loss
    uint32_t batch_cap = 0;
    uint64_t adam_t = 0;
    std::vector<rast_renderer *> users;
};

namespace rast_detail {

int ensure(DBuf &b, size_t bytes)
{
    if (b.p && b.n >= bytes) return RAST_OK;
    dfree(b);
    HIP_TRY(hipMalloc(&b.p, bytes < 256 ? 256 : bytes));
    b.n = bytes < 256 ? 256 : bytes;
    return RAST_OK;
}

int launch_infer(int H, uint32_t n, const float *x, const float *p, float margin, uint32_t slack, uint32_t *cap,
                 float *pred, hipStream_t s)
{
    const dim3 g(cdiv(n, 256)), b(256);
    switch (H) {
    case 16: hipLaunchKernelGGL(k_mlp_infer<16>, g, b, 0, s, n, x, p, margin, slack, cap, pred); break;
    case 32: hipLaunchKernelGGL(k_mlp_infer<32>, g, b, 0, s, n, x, p, margin, slack, cap, pred); break;
    case 64: hipLaunchKernelGGL(k_mlp_infer<64>, g, b, 0, s, n, x, p, margin, slack, cap, pred); break;
    default: return rast_set_error(RAST_E_INVALID_ARG, "hidden size %d not in {16, 32, 64}", H);
    }
    HIP_TRY(hipGetLastError());
    return RAST_OK;
}

constexpr uint32_t kMaxBatch = 1u << 16;

int ensure_batch(rast_binmodel *m, uint32_t B)
{
    if (B <= m->batch_cap) return RAST_OK;
    const size_t nb = (B + 63) / 64;
    RAST_TRY(ensure(m->idx, (size_t)B * 4));
    RAST_TRY(ensure(m->partial, nb * (num_params(m->H) + 1) * 4));
    RAST_TRY(ensure(m->loss, 64));
    m->batch_cap = B;
    return RAST_OK;
}

// One forward/backward over B samples (X, Y, optional index list) -> grad, loss.
int loss_grad(rast_binmodel *m, uint32_t B, uint32_t N, const float *X, const float *Y, const uint32_t *idx, float tau)
{
    if (B == 0 || B > kMaxBatch) return rast_set_error(RAST_E_INVALID_ARG, "batch %u not in [1, %u]", B, kMaxBatch);
    RAST_TRY(ensure_batch(m, B));
    const uint32_t nblocks = (B + 63) / 64;
    const uint32_t np = (uint32_t)num_params(m->H);
    const float *p = (const float *)m->params.p;
    float *part = (float *)m->partial.p;
    switch (m->H) {
    case 16: hipLaunchKernelGGL(k_train_partial<16>, dim3(nblocks), dim3(256), 0, m->stream, B, N, X, Y, idx, p, tau, part); break;
    case 32: hipLaunchKernelGGL(k_train_partial<32>, dim3(nblocks), dim3(256), 0, m->stream, B, N, X, Y, idx, p, tau, part); break;
    default: hipLaunchKernelGGL(k_train_partial<64>, dim3(nblocks), dim3(256), 0, m->stream, B, N, X, Y, idx, p, tau, part); break;
    }
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_partials, dim3(cdiv(np + 1, 256)), dim3(256), 0, m->stream, nblocks, np, B, part,
                       (float *)m->grad.p, (float *)m->loss.p);
    HIP_TRY(hipGetLastError());
    return RAST_OK;
}

}  // namespace rast_detail

using namespace rast_detail;

namespace rast {
hipError_t launch_tile_features(const FrameParams &fp, const uint32_t *prev_count, const float *prev_depth,
                                float *features, hipStream_t s)
{
    const uint32_t n = fp.tiles_x * fp.tiles_y;
    hipLaunchKernelGGL(k_features, dim3(cdiv(n, 256)), dim3(256), 0, s, fp, prev_count, prev_depth, features);
    return hipGetLastError();
}
}  // namespace rast


void rast_model_detach_renderer(rast_binmodel *m, rast_renderer *r)
{
    if (!m) return;
    for (auto &u : m->users)
        if (u == r) u = nullptr;
}

// --------------------------------------------------------------------- ABI

extern "C" int rast_binmodel_num_features(void) { return F; }

static void init_params(std::vector<float> &p, int H, uint32_t seed)
{
    uint32_t state = seed * 2654435761u + 12345u;
    auto next = [&]() {
        state ^= state << 13; state ^= state >> 17; state ^= state << 5;
        return (float)(state >> 8) * (1.0f / 16777216.0f);
    };
    size_t o = 0;
    const float a1 = std::sqrt(6.0f / F), a2 = std::sqrt(6.0f / H), a3 = std::sqrt(3.0f / H);
    for (int i = 0; i < H * F; ++i) p[o++] = (2.f * next() - 1.f) * a1;
    for (int i = 0; i < H; ++i) p[o++] = 0.01f;
    for (int i = 0; i < H * H; ++i) p[o++] = (2.f * next() - 1.f) * a2;
    for (int i = 0; i < H; ++i) p[o++] = 0.01f;
    for (int i = 0; i < H; ++i) p[o++] = (2.f * next() - 1.f) * a3;
    p[o++] = 1.0f;
}

extern "C" int rast_binmodel_create(rast_renderer *r, uint32_t hidden, uint32_t seed, rast_binmodel **out)
{
    if (!out) return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_create: null output");
    *out = nullptr;
    if (!r) return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_create: null renderer");
    if (hidden != 16 && hidden != 32 && hidden != 64)
        return rast_set_error(RAST_E_INVALID_ARG, "hidden size %u not in {16, 32, 64}", hidden);
    HIP_TRY(hipSetDevice(r->device));
    rast_binmodel *m = new (std::nothrow) rast_binmodel();
    if (!m) return rast_set_error(RAST_E_OUT_OF_MEMORY, "host allocation failed");
    m->device = r->device;
    m->H = (int)hidden;
    m->stream = r->stream;
    const size_t np = num_params(m->H);
    std::vector<float> p(np);
    init_params(p, m->H, seed);
    int rc = ensure(m->params, np * 4);
    if (rc == RAST_OK) rc = ensure(m->grad, np * 4);
    if (rc == RAST_OK) rc = ensure(m->adam_m, np * 4);
    if (rc == RAST_OK) rc = ensure(m->adam_v, np * 4);
    if (rc == RAST_OK && hipMemcpy(m->params.p, p.data(), np * 4, hipMemcpyHostToDevice) != hipSuccess) rc = RAST_E_HIP;
    if (rc == RAST_OK && hipMemset(m->adam_m.p, 0, np * 4) != hipSuccess) rc = RAST_E_HIP;
    if (rc == RAST_OK && hipMemset(m->adam_v.p, 0, np * 4) != hipSuccess) rc = RAST_E_HIP;
    if (rc != RAST_OK) {
        rast_binmodel_destroy(m);
        return rast_set_error(rc, "rast_binmodel_create: allocation failed");
    }
    m->users.push_back(r);
    *out = m;
    return RAST_OK;
}

extern "C" int rast_binmodel_destroy(rast_binmodel *m)
{
    if (!m) return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_destroy: null model");
    (void)hipSetDevice(m->device);
    for (rast_renderer *r : m->users)
        if (r) {
            (void)rast_settle(r);
            if (r->model == m) r->model = nullptr;
        }
    DBuf *bufs[] = { &m->params, &m->grad, &m->adam_m, &m->adam_v, &m->X, &m->Y, &m->idx, &m->partial, &m->loss,
                     &m->io_x, &m->io_y, &m->io_p };
    for (DBuf *b : bufs) dfree(*b);
    delete m;
    return RAST_OK;
}

extern "C" int rast_binmodel_num_params(rast_binmodel *m, size_t *out)
{
    if (!m || !out) return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_num_params: null argument");
    *out = num_params(m->H);
    return RAST_OK;
}

extern "C" int rast_binmodel_get_params(rast_binmodel *m, float *dst, size_t count)
{
    if (!m || !dst || count != num_params(m->H))
        return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_get_params: bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(dst, m->params.p, count * 4, hipMemcpyDeviceToHost));
    return RAST_OK;
}

extern "C" int rast_binmodel_set_params(rast_binmodel *m, const float *src, size_t count)
{
    if (!m || !src || count != num_params(m->H))
        return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_set_params: bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(m->params.p, src, count * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(m->adam_m.p, 0, count * 4));
    HIP_TRY(hipMemset(m->adam_v.p, 0, count * 4));
    m->adam_t = 0;
    return RAST_OK;
}

extern "C" int rast_binmodel_clear_data(rast_binmodel *m)
{
    if (!m) return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_clear_data: null model");
    m->n = 0;
    return RAST_OK;
}

extern "C" int rast_binmodel_record(rast_binmodel *m, rast_renderer *r)
{
    if (!m || !r) return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_record: null argument");
    RAST_TRY(rast_sync(r));
    if (!r->frame_valid || !r->frame_had_prev || r->model != m)
        return rast_set_error(RAST_E_STATE,
                              "rast_binmodel_record: needs two consecutive frames rendered with this model attached");
    HIP_TRY(hipSetDevice(m->device));
    const uint32_t tx = (r->W + r->tile - 1) / r->tile, ty = (r->H + r->tile - 1) / r->tile, nt = tx * ty;
    if ((uint64_t)m->n + nt > 0x7FFFFFFFull) return rast_set_error(RAST_E_STATE, "training set full");
    if (m->n + nt > m->cap) {
        const uint32_t ncap = std::max<uint32_t>(m->cap * 2, m->n + nt);
        DBuf nx, ny;
        if (hipMalloc(&nx.p, (size_t)ncap * F * 4) != hipSuccess) return rast_set_error(RAST_E_OUT_OF_MEMORY, "training set");
        nx.n = (size_t)ncap * F * 4;
        if (hipMalloc(&ny.p, (size_t)ncap * 4) != hipSuccess) {
            dfree(nx);
            return rast_set_error(RAST_E_OUT_OF_MEMORY, "training set");
        }
        ny.n = (size_t)ncap * 4;
        if (m->n) {
            HIP_TRY(hipMemcpy(nx.p, m->X.p, (size_t)m->n * F * 4, hipMemcpyDeviceToDevice));
            HIP_TRY(hipMemcpy(ny.p, m->Y.p, (size_t)m->n * 4, hipMemcpyDeviceToDevice));
        }
        dfree(m->X);
        dfree(m->Y);
        m->X = nx;
        m->Y = ny;
        m->cap = ncap;
    }
    // features of the frame just rendered (the previous frame's counts and
    // depths); target: this frame's exact counts
    FrameParams fp{};
    fp.tiles_x = tx;
    fp.tiles_y = ty;
    const int last = r->cur ^ 1, before = r->cur;
    float *xdst = (float *)m->X.p + (size_t)m->n * F;
    float *ydst = (float *)m->Y.p + (size_t)m->n;
    HIP_TRY(launch_tile_features(fp, (const uint32_t *)r->tile_count[before].p, (const float *)r->tile_min_depth[before].p, xdst,
                                 r->stream));
    hipLaunchKernelGGL(k_targets, dim3(cdiv(nt, 256)), dim3(256), 0, r->stream, nt,
                       (const uint32_t *)r->tile_count[last].p, ydst);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(r->stream));
    m->n += nt;
    return RAST_OK;
}

extern "C" int rast_binmodel_train(rast_binmodel *m, uint32_t epochs, uint32_t batch, float lr, float tau,
                                   float *final_loss)
{
    if (!m || batch == 0 || batch > kMaxBatch || !(lr > 0.f) || !(tau > 0.f && tau < 1.f))
        return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_train: bad arguments");
    if (m->n == 0) return rast_set_error(RAST_E_STATE, "rast_binmodel_train: no recorded data");
    HIP_TRY(hipSetDevice(m->device));
    RAST_TRY(ensure_batch(m, batch));
    const uint32_t steps = std::max<uint32_t>(1, (uint32_t)(((uint64_t)m->n * epochs + batch - 1) / batch));
    const uint32_t np = (uint32_t)num_params(m->H);
    for (uint32_t st = 0; st < steps; ++st) {
        hipLaunchKernelGGL(k_batch_index, dim3(cdiv(batch, 256)), dim3(256), 0, m->stream, batch, m->n,
                           (uint32_t)(m->adam_t * 0x9E3779B1ull + 7u), (uint32_t *)m->idx.p);
        RAST_TRY(loss_grad(m, batch, m->n, (const float *)m->X.p, (const float *)m->Y.p, (const uint32_t *)m->idx.p, tau));
        ++m->adam_t;
        const float b1c = 1.f - std::pow(0.9f, (float)m->adam_t), b2c = 1.f - std::pow(0.999f, (float)m->adam_t);
        hipLaunchKernelGGL(k_adam, dim3(cdiv(np, 256)), dim3(256), 0, m->stream, np, (float *)m->params.p,
                           (const float *)m->grad.p, (float *)m->adam_m.p, (float *)m->adam_v.p, lr, b1c, b2c);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(m->stream));
    if (final_loss) HIP_TRY(hipMemcpy(final_loss, m->loss.p, 4, hipMemcpyDeviceToHost));
    return RAST_OK;
}

extern "C" int rast_binmodel_forward(rast_binmodel *m, const float *x, uint32_t n, float *pred)
{
    if (!m || !x || !pred || n == 0) return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_forward: bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    RAST_TRY(ensure(m->io_x, (size_t)n * F * 4));
    RAST_TRY(ensure(m->io_p, (size_t)n * 4));
    HIP_TRY(hipMemcpy(m->io_x.p, x, (size_t)n * F * 4, hipMemcpyHostToDevice));
    RAST_TRY(launch_infer(m->H, n, (const float *)m->io_x.p, (const float *)m->params.p, 1.f, 0, nullptr,
                          (float *)m->io_p.p, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(pred, m->io_p.p, (size_t)n * 4, hipMemcpyDeviceToHost));
    return RAST_OK;
}

// ------------------------------------------------ allocation-free forward

// Device staging of rast_binmodel_forward_noalloc: at most
// kNoAllocMaxSamples samples (F floats each) and their predictions, 2.25 MiB
// on every GPU the library runs on, held until the process ends.  In host
// code this name is not a device address: the array's address on a device is
// read once with hipGetSymbolAddress (that device current) and used instead.
constexpr uint32_t kNoAllocMaxSamples = 1u << 16;
static __device__ __align__(256) float scratch[(size_t)kNoAllocMaxSamples * (kNumFeatures + 1)];

namespace {

constexpr int kNoAllocDevices = 64;

// Per-device address of `scratch`, and the lock that keeps calls on one GPU
// from overlapping (they share that GPU's array).
struct NoAllocState {
    std::mutex mu;
    float *ptr[kNoAllocDevices] = {};
};
NoAllocState g_noalloc;

// The calling thread holds g_noalloc.mu and has made `device` current.
int noalloc_scratch(int device, float **out)
{
    if (device < 0 || device >= kNoAllocDevices)
        return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_forward_noalloc: device %d above %d", device,
                              kNoAllocDevices - 1);
    if (!g_noalloc.ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return rast_set_error(RAST_E_HIP, "rast_binmodel_forward_noalloc: hipGetSymbolAddress(scratch): %s",
                                  hipGetErrorString(e));
        g_noalloc.ptr[device] = (float *)p;
    }
    *out = g_noalloc.ptr[device];
    return RAST_OK;
}

}  // namespace

extern "C" int rast_binmodel_forward_noalloc(rast_binmodel *m, const float *x, uint32_t n, float *pred)
{
    if (!m || !x || !pred || n == 0)
        return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_forward_noalloc: bad arguments");
    if (n > kNoAllocMaxSamples)
        return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_forward_noalloc: %u samples above the maximum %u", n,
                              kNoAllocMaxSamples);
    std::lock_guard<std::mutex> lock(g_noalloc.mu);
    HIP_TRY(hipSetDevice(m->device));
    float *base = nullptr;
    RAST_TRY(noalloc_scratch(m->device, &base));
    float *sx = base, *sp = base + (size_t)kNoAllocMaxSamples * F;
    // every word the kernel reads is written first: nothing to clear
    HIP_TRY(hipMemcpyAsync(sx, x, (size_t)n * F * 4, hipMemcpyHostToDevice, m->stream));
    RAST_TRY(launch_infer(m->H, n, sx, (const float *)m->params.p, 1.f, 0, nullptr, sp, m->stream));
    HIP_TRY(hipMemcpyAsync(pred, sp, (size_t)n * 4, hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return RAST_OK;
}

extern "C" int rast_binmodel_loss_grad(rast_binmodel *m, const float *x, const float *target, uint32_t n, float tau,
                                       float *loss, float *grad)
{
    if (!m || !x || !target || !grad || n == 0 || n > kMaxBatch)
        return rast_set_error(RAST_E_INVALID_ARG, "rast_binmodel_loss_grad: bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    RAST_TRY(ensure(m->io_x, (size_t)n * F * 4));
    RAST_TRY(ensure(m->io_y, (size_t)n * 4));
    HIP_TRY(hipMemcpy(m->io_x.p, x, (size_t)n * F * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->io_y.p, target, (size_t)n * 4, hipMemcpyHostToDevice));
    RAST_TRY(loss_grad(m, n, n, (const float *)m->io_x.p, (const float *)m->io_y.p, nullptr, tau));
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(grad, m->grad.p, num_params(m->H) * 4, hipMemcpyDeviceToHost));
    if (loss) HIP_TRY(hipMemcpy(loss, m->loss.p, 4, hipMemcpyDeviceToHost));
    return RAST_OK;
}

int rast_model_hidden(const rast_binmodel *m) { return m ? m->H : 0; }
const float *rast_model_params(const rast_binmodel *m) { return m ? (const float *)m->params.p : nullptr; }
