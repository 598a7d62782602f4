// rast_api.hip -- host side of librast.so: the C ABI of include/rast.h.
//
// Owns the device objects (renderer, meshes, textures), computes the frame
// constants with the reference's host camera math, and enqueues one launch
// per pipeline stage on the renderer's stream.  Every HIP status is checked
// and mapped to a RAST_E_* code; nothing throws across the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "rast.h"
#include "rast_launch.h"
#include "rast_model.h"
#include "rast_state.h"

using namespace rast;

// ----------------------------------------------------------------- errors

static thread_local std::string g_last_error;

int rast_set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

extern "C" int rast_last_error(char *buf, size_t n)
{
    const size_t len = g_last_error.size();
    if (buf && n) {
        const size_t c = len < n - 1 ? len : n - 1;
        memcpy(buf, g_last_error.data(), c);
        buf[c] = 0;
    }
    return (int)len;
}

// ------------------------------------------------------------ host math

// Matrices use the reference storage data[col][row] (CCR/Matrix.h:9-44).
static inline float &M4(float *m, int r, int c) { return m[c * 4 + r]; }
static inline float M4c(const float *m, int r, int c) { return m[c * 4 + r]; }

// Matrix4 * Matrix4 (CCR/Matrix4.h:120-148), sums left to right.
static void mat4_mul(const float *lm, const float *rm, float *out)
{
    float t[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
            M4(t, r, c) = M4c(lm, r, 0) * M4c(rm, 0, c) + M4c(lm, r, 1) * M4c(rm, 1, c) +
                          M4c(lm, r, 2) * M4c(rm, 2, c) + M4c(lm, r, 3) * M4c(rm, 3, c);
    memcpy(out, t, sizeof t);
}

struct V3f { float x, y, z; };
static V3f cross3(V3f a, V3f b) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
static float dot3(V3f a, V3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V3f normalized(V3f v)
{  // CCR/Vector.h:41-63
    const float m = std::sqrt(dot3(v, v));
    if (m == 0.f) return { 0.f, 0.f, 0.f };
    const float inv = 1.f / m;
    return { v.x * inv, v.y * inv, v.z * inv };
}

// Inverse (CCR/Matrix4.h:327-363), FGED1 formulation in the same order.
static void mat4_inverse(const float *m, float *out)
{
    const V3f a{ m[0], m[1], m[2] }, b{ m[4], m[5], m[6] }, c{ m[8], m[9], m[10] }, d{ m[12], m[13], m[14] };
    const float x = M4c(m, 3, 0), y = M4c(m, 3, 1), z = M4c(m, 3, 2), w = M4c(m, 3, 3);
    V3f s = cross3(a, b), t = cross3(c, d);
    V3f u{ a.x * y - b.x * x, a.y * y - b.y * x, a.z * y - b.z * x };
    V3f v{ c.x * w - d.x * z, c.y * w - d.y * z, c.z * w - d.z * z };
    const float det = dot3(s, v) + dot3(t, u);
    const float invDet = 1.0f / det;
    s = { s.x * invDet, s.y * invDet, s.z * invDet };
    t = { t.x * invDet, t.y * invDet, t.z * invDet };
    u = { u.x * invDet, u.y * invDet, u.z * invDet };
    v = { v.x * invDet, v.y * invDet, v.z * invDet };
    const V3f bv = cross3(b, v), va = cross3(v, a), du = cross3(d, u), uc = cross3(u, c);
    const V3f r0{ bv.x + t.x * y, bv.y + t.y * y, bv.z + t.z * y };
    const V3f r1{ va.x - t.x * x, va.y - t.y * x, va.z - t.z * x };
    const V3f r2{ du.x + s.x * w, du.y + s.y * w, du.z + s.z * w };
    const V3f r3{ uc.x - s.x * z, uc.y - s.y * z, uc.z - s.z * z };
    const float rows[4][4] = { { r0.x, r0.y, r0.z, -dot3(b, t) },
                               { r1.x, r1.y, r1.z, dot3(a, t) },
                               { r2.x, r2.y, r2.z, -dot3(d, s) },
                               { r3.x, r3.y, r3.z, dot3(c, s) } };
    for (int r = 0; r < 4; ++r)
        for (int cc = 0; cc < 4; ++cc) M4(out, r, cc) = rows[r][cc];
}

// Camera (CCR/Camera.cpp:6-31,97-117): look-at = {right, up, forward, pos};
// view = Inverse(look-at); projection keeps Identity()'s [3][3] = 1.
static void camera_frame(const rast_camera &cam, uint32_t W, uint32_t H, rast_frame_consts &fr)
{
    const V3f worldUp{ 0.f, 1.f, 0.f };
    const V3f fwd = normalized({ cam.forward[0], cam.forward[1], cam.forward[2] });
    const V3f right = normalized(cross3(worldUp, fwd));
    const V3f up = normalized(cross3(fwd, right));
    float lookAt[16] = { right.x, right.y, right.z, 0.f, up.x, up.y, up.z, 0.f,
                         fwd.x, fwd.y, fwd.z, 0.f, cam.position[0], cam.position[1], cam.position[2], 1.f };
    const float pi_f = 3.14159265358979323846f;
    const float radians = cam.fov_deg * (pi_f / 180.f);
    const float fov = std::tan(radians / 2.f);
    const float aspect = (float)W / (float)H;
    const float n = cam.near_z, f = cam.far_z;
    const float A = f / (n - f);
    const float B = (f * n) / (n - f);
    float *P = fr.proj;
    for (int i = 0; i < 16; ++i) P[i] = (i % 5 == 0) ? 1.f : 0.f;
    P[0] = 1.f / (aspect * fov);
    P[5] = 1.f / fov;
    P[10] = A;
    P[11] = -1.f;  // data[2][3]
    P[14] = B;     // data[3][2]
    mat4_inverse(lookAt, fr.view);
    mat4_mul(fr.proj, fr.view, fr.viewproj);
    const float cp[3] = { cam.position[0], cam.position[1], cam.position[2] };
    const float cf[3] = { fwd.x, fwd.y, fwd.z }, cr[3] = { right.x, right.y, right.z }, cu[3] = { up.x, up.y, up.z };
    memcpy(fr.cam_pos, cp, sizeof cp);
    memcpy(fr.cam_fwd, cf, sizeof cf);
    memcpy(fr.cam_right, cr, sizeof cr);
    memcpy(fr.cam_up, cu, sizeof cu);
}

// Batches run on one process-wide stream.  A renderer whose last frame came
// from batch k waits on the batch event (re-recorded by every batch, so it
// covers k) unless a batch at least as recent is already known complete.
struct BatchSync {
    std::atomic<uint64_t> recorded{ 0 };
    std::atomic<uint64_t> waited{ 0 };
    hipEvent_t done = nullptr;
};
static BatchSync g_bsync;

static hipError_t renderer_wait(rast_renderer *r)
{
    if (r->batch_seq) {
        if (r->batch_seq > g_bsync.waited.load()) {
            const uint64_t target = g_bsync.recorded.load();
            const hipError_t e = hipEventSynchronize(g_bsync.done);
            if (e != hipSuccess) return e;
            uint64_t w = g_bsync.waited.load();
            while (w < target && !g_bsync.waited.compare_exchange_weak(w, target)) {
            }
        }
        r->batch_seq = 0;
    }
    if (r->stream_busy) {
        const hipError_t e = hipStreamSynchronize(r->stream);
        if (e != hipSuccess) return e;
        r->stream_busy = false;
    }
    return hipSuccess;
}

int rast_settle(rast_renderer *r)
{
    HIP_TRY(renderer_wait(r));
    return RAST_OK;
}

static hipError_t grow(rast_renderer *r, DBuf &b, size_t bytes, bool *synced)
{
    if (b.p && b.n >= bytes) return hipSuccess;
    if (!*synced) {
        hipError_t e = renderer_wait(r);
        if (e != hipSuccess) return e;
        *synced = true;
    }
    dfree(b);
    size_t nb = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&b.p, nb);
    if (e != hipSuccess) { b.p = nullptr; return e; }
    b.n = nb;
    return hipSuccess;
}

static uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

static uint32_t mesh_triangles(int topology, uint32_t ni)
{  // CCR/CUDAMesh.cpp:27-40
    return topology == RAST_TOPOLOGY_LIST ? ni / 3 : (ni >= 3 ? ni - 2 : 0);
}

extern "C" int rast_device_query(int device, rast_device_info *out)
{
    if (!out) return rast_set_error(RAST_E_INVALID_ARG, "rast_device_query: null output");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return rast_set_error(RAST_E_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= n) return rast_set_error(RAST_E_INVALID_ARG, "device %d out of range", device);
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, device));
    memset(out, 0, sizeof *out);
    snprintf(out->name, sizeof out->name, "%s", p.name);
    snprintf(out->arch, sizeof out->arch, "%s", p.gcnArchName);
    out->compute_units = p.multiProcessorCount;
    out->clock_khz = p.clockRate;
    out->global_mem_bytes = p.totalGlobalMem;
    out->lds_bytes_per_block = (int)p.sharedMemPerBlock;
    out->wavefront_size = p.warpSize;
    return RAST_OK;
}

extern "C" int rast_camera_frame(const rast_camera *cam, uint32_t width, uint32_t height, rast_frame_consts *out)
{
    if (!cam || !out || !width || !height) return rast_set_error(RAST_E_INVALID_ARG, "rast_camera_frame: bad arguments");
    camera_frame(*cam, width, height, *out);
    return RAST_OK;
}

// Timing events of one frame, recorded between the stages when RAST_FLAG_TIMING is set.
enum { EV_START, EV_VERTEX, EV_MODEL, EV_SETUP, EV_SCANCNT, EV_FILL, EV_FALLBACK, EV_RASTER, EV_SHADE, EV_COUNT };
// the start / stop events of rast_timer_start / rast_timer_stop
enum { EV_TIMER_START = EV_COUNT, EV_TIMER_STOP, EV_ALL };
static_assert(EV_ALL <= (int)(sizeof(((rast_renderer *)nullptr)->ev) / sizeof(hipEvent_t)), "renderer event slots");

static int alloc_framebuffers(rast_renderer *r)
{
    bool synced = false;
    const size_t P = (size_t)r->W * r->H;
    HIP_TRY(grow(r, r->vis, P * 4, &synced));
    HIP_TRY(grow(r, r->depth, P * 4, &synced));
    HIP_TRY(grow(r, r->rgba, P * 4, &synced));
    HIP_TRY(grow(r, r->keys, P * 8, &synced));
    r->keys_dirty = true;
    r->frame_valid = false;
    return RAST_OK;
}

static int new_renderer(int device, uint32_t width, uint32_t height, rast_renderer *scene_src, rast_renderer **out)
{
    if (!out) return rast_set_error(RAST_E_INVALID_ARG, "rast_renderer_create: null output");
    *out = nullptr;
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return rast_set_error(RAST_E_INVALID_ARG, "framebuffer %ux%u out of range (1..16384)", width, height);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return rast_set_error(RAST_E_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= n) return rast_set_error(RAST_E_INVALID_ARG, "device %d out of range", device);
    HIP_TRY(hipSetDevice(device));
    rast_renderer *r = new (std::nothrow) rast_renderer();
    if (!r) return rast_set_error(RAST_E_OUT_OF_MEMORY, "host allocation failed");
    r->device = device;
    r->W = width;
    r->H = height;
    int rc = RAST_OK;
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) rc = RAST_E_HIP;
    if (rc == RAST_OK && hipEventCreateWithFlags(&r->join, hipEventDisableTiming) != hipSuccess) rc = RAST_E_HIP;
    for (int i = 0; rc == RAST_OK && i < EV_ALL; ++i)
        if (hipEventCreate(&r->ev[i]) != hipSuccess) rc = RAST_E_HIP;
    if (rc == RAST_OK && hipHostMalloc((void **)&r->h_counters, sizeof(FrameCounters)) != hipSuccess) rc = RAST_E_HIP;
    if (rc == RAST_OK && hipHostGetDevicePointer((void **)&r->d_counters_host, r->h_counters, 0) != hipSuccess)
        rc = RAST_E_HIP;
    if (rc == RAST_OK) rc = alloc_framebuffers(r);
    if (rc != RAST_OK) {
        rast_renderer_destroy(r);
        return rast_set_error(rc, "rast_renderer_create: HIP resource allocation failed");
    }
    memset(r->h_counters, 0, sizeof(FrameCounters));
    if (scene_src) {
        r->scene_src = scene_src;
        scene_src->views.push_back(r);
    }
    *out = r;
    return RAST_OK;
}

extern "C" int rast_renderer_create(int device, uint32_t width, uint32_t height, rast_renderer **out)
{
    return new_renderer(device, width, height, nullptr, out);
}

extern "C" int rast_renderer_create_view(rast_renderer *owner, uint32_t width, uint32_t height, rast_renderer **out)
{
    if (!owner) return rast_set_error(RAST_E_INVALID_ARG, "rast_renderer_create_view: null owner");
    if (owner->scene_src) owner = owner->scene_src;  // a view of a view draws the same scene
    return new_renderer(owner->device, width, height, owner, out);
}

static void free_mesh(rast_mesh *m)
{
    if (m->ivert14) (void)hipFree(m->ivert14);
    if (m->idx) (void)hipFree(m->idx);
    delete m;
}

// Frees a mesh's device memory and detaches it from its renderer; the handle
// stays allocated until rast_mesh_destroy.
static void orphan_mesh(rast_mesh *m)
{
    if (m->ivert14) (void)hipFree(m->ivert14);
    if (m->idx) (void)hipFree(m->idx);
    m->ivert14 = nullptr;
    m->idx = nullptr;
    for (auto &t : m->tex) t = nullptr;
    m->texq = nullptr;
    m->r = nullptr;
}

extern "C" int rast_renderer_destroy(rast_renderer *r)
{
    if (!r) return rast_set_error(RAST_E_INVALID_ARG, "rast_renderer_destroy: null renderer");
    if (!r->views.empty())
        return rast_set_error(RAST_E_STATE, "rast_renderer_destroy: %zu views still draw this renderer's meshes",
                              r->views.size());
    (void)hipSetDevice(r->device);
    if (r->stream) (void)renderer_wait(r);
    if (r->scene_src) {
        std::vector<rast_renderer *> &v = r->scene_src->views;
        v.erase(std::remove(v.begin(), v.end(), r), v.end());
    }
    plan_free(r->plan);
    rast_model_detach_renderer(r->model, r);
    // Mesh and texture handles outlive the renderer as empty handles: their
    // device memory goes now, the handle itself with rast_mesh_destroy /
    // rast_texture_destroy, so a caller may destroy the renderer before the
    // objects that hold those handles (the reference app deletes its
    // CUDARenderer before its CUDATextureManager, CCR/main.cpp:418,529).
    for (rast_mesh *m : r->meshes) orphan_mesh(m);
    for (TexQuad &q : r->tex_quads) (void)hipFree(q.texels);
    for (rast_texture *t : r->textures) {
        if (t->texels) (void)hipFree(t->texels);
        t->texels = nullptr;
        t->r = nullptr;
    }
    DBuf *bufs[] = { &r->pos4, &r->attr, &r->trirec,
                     &r->alive, &r->survivors, &r->tile_count[0], &r->tile_count[1], &r->tile_off, &r->tile_cap,
                     &r->tile_pool, &r->tile_cursor, &r->tile_min_depth[0], &r->tile_min_depth[1], &r->keys,
                     &r->queue, &r->pool, &r->big, &r->bigp, &r->ovf, &r->vis, &r->depth,
                     &r->rgba };
    for (DBuf *b : bufs) dfree(*b);
    if (r->h_counters) (void)hipHostFree(r->h_counters);
    for (hipEvent_t &e : r->ev)
        if (e) (void)hipEventDestroy(e);
    if (r->join) (void)hipEventDestroy(r->join);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
    return RAST_OK;
}

extern "C" int rast_setup_bins(rast_renderer *r, uint32_t tile_size, uint32_t big_threshold)
{
    if (!r) return rast_set_error(RAST_E_INVALID_ARG, "rast_setup_bins: null renderer");
    if (tile_size != 8 && tile_size != 16 && tile_size != 32)
        return rast_set_error(RAST_E_INVALID_ARG, "tile size %u not in {8, 16, 32}", tile_size);
    r->tile = tile_size;
    r->big_threshold = big_threshold ? big_threshold : 256;
    r->have_prev = false;
    return RAST_OK;
}

extern "C" int rast_set_tuning(rast_renderer *r, uint32_t micro_limit, uint32_t small_limit)
{
    if (!r) return rast_set_error(RAST_E_INVALID_ARG, "rast_set_tuning: null renderer");
    if (micro_limit > 1024 || small_limit > 1024)
        return rast_set_error(RAST_E_INVALID_ARG, "micro_limit %u / small_limit %u above 1024", micro_limit, small_limit);
    r->micro_limit = micro_limit;
    r->small_limit = small_limit;
    return RAST_OK;
}

extern "C" int rast_set_resolve_mode(rast_renderer *r, int mode)
{
    if (!r) return rast_set_error(RAST_E_INVALID_ARG, "rast_set_resolve_mode: null renderer");
    if (mode < RAST_RESOLVE_PER_TILE || mode > RAST_RESOLVE_PERSISTENT_SHADE)
        return rast_set_error(RAST_E_INVALID_ARG, "resolve mode %d not in [0, 2]", mode);
    r->resolve_mode = mode;
    return RAST_OK;
}

extern "C" int rast_resize(rast_renderer *r, uint32_t width, uint32_t height)
{
    if (!r) return rast_set_error(RAST_E_INVALID_ARG, "rast_resize: null renderer");
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return rast_set_error(RAST_E_INVALID_ARG, "framebuffer %ux%u out of range (1..16384)", width, height);
    HIP_TRY(hipSetDevice(r->device));
    r->W = width;
    r->H = height;
    r->have_prev = false;
    return alloc_framebuffers(r);
}

// ------------------------------------------------------------------ meshes

extern "C" int rast_mesh_create(rast_renderer *r, const rast_ivertex *verts, uint32_t num_verts,
                                const uint32_t *indices, uint32_t num_indices, int topology, const float world[16],
                                rast_mesh **out)
{
    if (!out) return rast_set_error(RAST_E_INVALID_ARG, "rast_mesh_create: null output");
    *out = nullptr;
    if (!r || !verts || !indices || num_verts == 0 || num_indices == 0)
        return rast_set_error(RAST_E_INVALID_ARG, "rast_mesh_create: empty or null mesh data");
    if (r && r->scene_src) return rast_set_error(RAST_E_STATE, "rast_mesh_create: a view has no meshes of its own");
    if (topology != RAST_TOPOLOGY_LIST && topology != RAST_TOPOLOGY_STRIP)
        return rast_set_error(RAST_E_INVALID_ARG, "unknown topology %d", topology);
    if (r->meshes.size() >= (size_t)kMaxMeshes) return rast_set_error(RAST_E_INVALID_ARG, "too many meshes");
    const uint32_t ntri = mesh_triangles(topology, num_indices);
    uint64_t tris = ntri, verts_total = num_verts;
    for (rast_mesh *m : r->meshes) { tris += m->ntri; verts_total += m->nv; }
    if (tris >= 0xFFFFFFF0ull || verts_total >= 0xFFFFFFF0ull)
        return rast_set_error(RAST_E_INVALID_ARG, "scene exceeds 2^32 triangles or vertices");
    for (uint32_t i = 0; i < num_indices; ++i)
        if (indices[i] >= num_verts)
            return rast_set_error(RAST_E_INVALID_ARG, "index %u at position %u out of range (num_verts %u)", indices[i], i,
                        num_verts);
    HIP_TRY(hipSetDevice(r->device));
    static_assert(sizeof(rast_ivertex) == 14 * sizeof(float), "IVertex: 14 floats");
    // as given, plus 16 bytes: the vertex pass reads whole 16-byte words
    std::vector<float> v14((size_t)num_verts * 14 + 4, 0.f);
    float attr_max = 0.f;
    for (uint32_t i = 0; i < num_verts; ++i) {
        memcpy(&v14[(size_t)i * 14], &verts[i], sizeof(rast_ivertex));
        for (int k = 0; k < 3; ++k) {
            const float a = std::fabs(verts[i].n[k]), b = std::fabs(verts[i].tan[k]);
            attr_max = std::max(attr_max, std::max(a, b));
            if (a != a || b != b) attr_max = INFINITY;
        }
    }
    rast_mesh *m = new (std::nothrow) rast_mesh();
    if (!m) return rast_set_error(RAST_E_OUT_OF_MEMORY, "host allocation failed");
    m->r = r;
    m->nv = num_verts;
    m->ni = num_indices;
    m->ntri = ntri;
    m->topology = topology;
    m->attr_max = attr_max;
    for (int i = 0; i < 16; ++i) m->world[i] = world ? world[i] : ((i % 5 == 0) ? 1.f : 0.f);
    hipError_t e = hipMalloc((void **)&m->ivert14, v14.size() * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&m->idx, (size_t)num_indices * 4);
    if (e == hipSuccess) e = hipMemcpy(m->ivert14, v14.data(), v14.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->idx, indices, (size_t)num_indices * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_mesh(m);
        return rast_set_error(e == hipErrorOutOfMemory ? RAST_E_OUT_OF_MEMORY : RAST_E_HIP, "mesh upload: %s",
                    hipGetErrorString(e));
    }
    r->meshes.push_back(m);
    ++r->scene_version;
    r->have_prev = false;
    *out = m;
    return RAST_OK;
}

extern "C" int rast_mesh_destroy(rast_mesh *m)
{
    if (!m) return rast_set_error(RAST_E_INVALID_ARG, "rast_mesh_destroy: null mesh");
    rast_renderer *r = m->r;
    if (!r) {  // its renderer is gone (and with it the device memory)
        delete m;
        return RAST_OK;
    }
    (void)hipSetDevice(r->device);
    HIP_TRY(renderer_wait(r));
    for (rast_renderer *v : r->views) {
        HIP_TRY(renderer_wait(v));
        v->have_prev = false;
    }
    r->meshes.erase(std::remove(r->meshes.begin(), r->meshes.end(), m), r->meshes.end());
    free_mesh(m);
    ++r->scene_version;
    r->have_prev = false;
    r->frame_valid = false;
    return RAST_OK;
}

extern "C" int rast_mesh_set_world(rast_mesh *m, const float world[16])
{
    if (!m || !world) return rast_set_error(RAST_E_INVALID_ARG, "rast_mesh_set_world: null argument");
    memcpy(m->world, world, sizeof m->world);
    return RAST_OK;
}

extern "C" int rast_mesh_set_textures(rast_mesh *m, rast_texture *d, rast_texture *n, rast_texture *s,
                                      rast_texture *g)
{
    if (!m) return rast_set_error(RAST_E_INVALID_ARG, "rast_mesh_set_textures: null mesh");
    if (!m->r) return rast_set_error(RAST_E_STATE, "rast_mesh_set_textures: the mesh's renderer was destroyed");
    rast_texture *t[4] = { d, n, s, g };
    for (int i = 0; i < 4; ++i)
        if (t[i] && t[i]->r != m->r) return rast_set_error(RAST_E_INVALID_ARG, "texture belongs to another renderer");
    for (int i = 0; i < 4; ++i) m->tex[i] = t[i];
    return RAST_OK;
}

extern "C" int rast_mesh_num_triangles(const rast_mesh *m, uint32_t *out)
{
    if (!m || !out) return rast_set_error(RAST_E_INVALID_ARG, "rast_mesh_num_triangles: null argument");
    *out = m->ntri;
    return RAST_OK;
}

extern "C" int rast_mesh_triangle_base(const rast_mesh *m, uint64_t *out)
{
    if (!m || !out) return rast_set_error(RAST_E_INVALID_ARG, "rast_mesh_triangle_base: null argument");
    if (!m->r) return rast_set_error(RAST_E_STATE, "rast_mesh_triangle_base: the mesh's renderer was destroyed");
    uint64_t base = 0;
    for (rast_mesh *q : m->r->meshes) {
        if (q == m) break;
        base += q->ntri;
    }
    *out = base;
    return RAST_OK;
}

// ---------------------------------------------------------------- textures

extern "C" int rast_texture_create_pitched(rast_renderer *r, const void *rgba8, uint32_t w, uint32_t h,
                                           size_t pitch_bytes, rast_texture **out)
{
    if (!out) return rast_set_error(RAST_E_INVALID_ARG, "rast_texture_create: null output");
    *out = nullptr;
    if (!r || !rgba8 || !w || !h || w > 32768 || h > 32768)
        return rast_set_error(RAST_E_INVALID_ARG, "rast_texture_create: bad texture %ux%u", w, h);
    if (pitch_bytes < (size_t)w * 4 || pitch_bytes % 4 != 0)
        return rast_set_error(RAST_E_INVALID_ARG, "rast_texture_create: pitch %zu below 4 x width or not a multiple of 4",
                              pitch_bytes);
    HIP_TRY(hipSetDevice(r->device));
    rast_texture *t = new (std::nothrow) rast_texture();
    if (!t) return rast_set_error(RAST_E_OUT_OF_MEMORY, "host allocation failed");
    t->r = r;
    t->w = w;
    t->h = h;
    const size_t row = (size_t)w * 4;
    // stored packed (pitch = width): the sampler addresses texel x + y * w
    hipError_t e = hipMalloc((void **)&t->texels, row * h);
    if (e == hipSuccess) e = hipMemcpy2D(t->texels, row, rgba8, pitch_bytes, row, h, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (t->texels) (void)hipFree(t->texels);
        delete t;
        return rast_set_error(e == hipErrorOutOfMemory ? RAST_E_OUT_OF_MEMORY : RAST_E_HIP, "texture upload: %s",
                    hipGetErrorString(e));
    }
    r->textures.push_back(t);
    *out = t;
    return RAST_OK;
}

extern "C" int rast_texture_create(rast_renderer *r, const uint32_t *rgba8, uint32_t w, uint32_t h,
                                   rast_texture **out)
{
    return rast_texture_create_pitched(r, rgba8, w, h, (size_t)w * 4, out);
}

extern "C" int rast_texture_destroy(rast_texture *t)
{
    if (!t) return rast_set_error(RAST_E_INVALID_ARG, "rast_texture_destroy: null texture");
    rast_renderer *r = t->r;
    if (!r) {  // its renderer is gone (and with it the texels)
        delete t;
        return RAST_OK;
    }
    (void)hipSetDevice(r->device);
    HIP_TRY(renderer_wait(r));
    for (rast_renderer *v : r->views) HIP_TRY(renderer_wait(v));
    for (rast_mesh *m : r->meshes) {
        for (auto &slot : m->tex)
            if (slot == t) slot = nullptr;
        m->texq = nullptr;  // set again by the next frame
    }
    for (size_t i = r->tex_quads.size(); i-- > 0;) {
        const TexQuad &q = r->tex_quads[i];
        if (q.t[0] == t || q.t[1] == t || q.t[2] == t || q.t[3] == t) {
            (void)hipFree(q.texels);
            r->tex_quads.erase(r->tex_quads.begin() + (std::ptrdiff_t)i);
        }
    }
    r->textures.erase(std::remove(r->textures.begin(), r->textures.end(), t), r->textures.end());
    if (t->texels) (void)hipFree(t->texels);
    delete t;
    return RAST_OK;
}

// ------------------------------------------------------------------ frame

// Renderers created with rast_renderer_create_view draw the meshes of their
// owner; everything per frame (buffers, launch plan) is their own.
static std::vector<rast_mesh *> &scene_meshes(rast_renderer *r) { return r->scene_src ? r->scene_src->meshes : r->meshes; }
static rast_renderer *scene_owner(rast_renderer *r) { return r->scene_src ? r->scene_src : r; }

static void refresh_layout(rast_renderer *r)
{
    uint64_t vb = 0, tb = 0;
    for (rast_mesh *m : scene_meshes(r)) {
        m->tri_base = tb;
        vb += m->nv;
        tb += m->ntri;
    }
    r->total_verts = (uint32_t)vb;
    r->total_tris = (uint32_t)tb;
    r->layout_version = scene_owner(r)->scene_version;
    r->have_prev = false;
}

static int ensure_frame_storage(rast_renderer *r, uint32_t ntiles)
{
    bool synced = false;
    const size_t nv = (size_t)r->total_verts + 1, nt = (size_t)r->total_tris + 1;
    HIP_TRY(grow(r, r->pos4, nv * 16, &synced));
    HIP_TRY(grow(r, r->attr, nv * 48, &synced));
    HIP_TRY(grow(r, r->trirec, nt * kTriRecWords * 4, &synced));
    HIP_TRY(grow(r, r->alive, nt, &synced));
    HIP_TRY(grow(r, r->survivors, nt * 4, &synced));
    const size_t tb = ((size_t)ntiles + 64) * 4;
    const void *prev_bufs[4] = { r->tile_count[0].p, r->tile_count[1].p, r->tile_min_depth[0].p,
                                 r->tile_min_depth[1].p };
    for (int i = 0; i < 2; ++i) {
        HIP_TRY(grow(r, r->tile_count[i], tb, &synced));
        HIP_TRY(grow(r, r->tile_min_depth[i], tb, &synced));
    }
    // the previous frame's per-tile counts and depths are gone if their buffers moved
    if (prev_bufs[0] != r->tile_count[0].p || prev_bufs[1] != r->tile_count[1].p ||
        prev_bufs[2] != r->tile_min_depth[0].p || prev_bufs[3] != r->tile_min_depth[1].p)
        r->have_prev = false;
    HIP_TRY(grow(r, r->tile_off, tb, &synced));
    HIP_TRY(grow(r, r->tile_cap, tb, &synced));
    HIP_TRY(grow(r, r->tile_pool, tb, &synced));
    HIP_TRY(grow(r, r->tile_cursor, tb, &synced));
    if (r->queue_cap == 0) r->queue_cap = std::max<uint32_t>(1u << 20, std::min<uint64_t>(4ull * nt, 0x7FFFFFFFull));
    if (r->pool_cap == 0) r->pool_cap = 1u << 20;
    if (r->big_cap == 0) r->big_cap = 1u << 16;
    if (r->ovf_cap == 0) r->ovf_cap = 1u << 20;
    HIP_TRY(grow(r, r->ovf, (size_t)r->ovf_cap * 8, &synced));
    HIP_TRY(grow(r, r->queue, (size_t)r->queue_cap * 4, &synced));
    HIP_TRY(grow(r, r->pool, (size_t)r->pool_cap * 4, &synced));
    HIP_TRY(grow(r, r->big, (size_t)r->big_cap * 4, &synced));
    HIP_TRY(grow(r, r->bigp, (size_t)r->big_cap * 16, &synced));
    return RAST_OK;
}

static int validate_opts(const rast_frame_opts *o)
{
    if (o->cull_mode < 0 || o->cull_mode > 2) return rast_set_error(RAST_E_INVALID_ARG, "cull_mode %d", o->cull_mode);
    if (o->sample_state < 0 || o->sample_state > 1)
        return rast_set_error(RAST_E_INVALID_ARG, "sample_state %d", o->sample_state);
    if (o->visualisation < 0 || o->visualisation > 2)
        return rast_set_error(RAST_E_INVALID_ARG, "visualisation %d", o->visualisation);
    if (o->bin_policy < 0 || o->bin_policy > 2) return rast_set_error(RAST_E_INVALID_ARG, "bin_policy %d", o->bin_policy);
    return RAST_OK;
}

static int validate_camera(const rast_camera *cam)
{
    if (!(cam->near_z > 0.f) || !(cam->far_z > cam->near_z) || !(cam->fov_deg > 0.f && cam->fov_deg < 180.f))
        return rast_set_error(RAST_E_INVALID_ARG, "camera near/far/fov out of range");
    return RAST_OK;
}

// Host part of a renderer's next frame (r->last_cam / last_opts): scene
// layout, storage and the binning mode.
// Points each mesh whose four maps exist and share one size at their
// interleaved copy, building it on first use (once per set of four
// textures; freed with any of them).  Sets above 2^28 texels keep the
// separate maps: shading addresses the copy with 32-bit byte offsets.
static int ensure_tex_quads(rast_renderer *r)
{
    rast_renderer *owner = scene_owner(r);
    for (rast_mesh *m : scene_meshes(r)) {
        m->texq = nullptr;
        rast_texture *const *t = m->tex;
        if (!t[0] || !t[1] || !t[2] || !t[3]) continue;
        const uint32_t w = t[0]->w, h = t[0]->h;
        bool same = true;
        for (int k = 1; k < 4; ++k) same = same && t[k]->w == w && t[k]->h == h;
        if (!same || (uint64_t)w * h > (1ull << 28)) continue;
        TexQuad *found = nullptr;
        for (TexQuad &q : owner->tex_quads)
            if (q.t[0] == t[0] && q.t[1] == t[1] && q.t[2] == t[2] && q.t[3] == t[3]) found = &q;
        if (!found) {
            TexQuad q = { { t[0], t[1], t[2], t[3] }, nullptr };
            const size_t n = (size_t)w * h;
            const hipError_t me = hipMalloc((void **)&q.texels, n * sizeof(uint4));
            if (me == hipErrorOutOfMemory) {
                // the copy is an optimisation only: without room for it the
                // frame samples the four separate maps
                (void)hipGetLastError();
                q.texels = nullptr;
                continue;
            }
            HIP_TRY(me);
            hipError_t e = launch_interleave_maps(t[0]->texels, t[1]->texels, t[2]->texels, t[3]->texels, q.texels, n,
                                                  r->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
            if (e != hipSuccess) {
                (void)hipFree(q.texels);
                return rast_set_error(RAST_E_HIP, "interleaving the maps: %s", hipGetErrorString(e));
            }
            owner->tex_quads.push_back(q);
            found = &owner->tex_quads.back();
        }
        m->texq = found->texels;
    }
    return RAST_OK;
}

static int prepare_job(rast_renderer *r)
{
    HIP_TRY(hipSetDevice(r->device));
    if (scene_meshes(r).empty()) return rast_set_error(RAST_E_STATE, "rast_render: no meshes loaded");
    if (r->layout_version != scene_owner(r)->scene_version) refresh_layout(r);
    RAST_TRY(ensure_tex_quads(r));
    RAST_TRY(ensure_frame_storage(r, ceil_div(r->W, r->tile) * ceil_div(r->H, r->tile)));
    r->frame_had_prev = r->have_prev;
    bool single = r->last_opts.bin_policy != RAST_BIN_EXACT && r->have_prev;
    if (r->last_opts.bin_policy == RAST_BIN_LEARNED && !r->model) single = false;
    r->last_single = single;
    return RAST_OK;
}

static bool job_learned(const rast_renderer *r) { return r->last_single && r->last_opts.bin_policy == RAST_BIN_LEARNED; }
static bool job_shades(const rast_renderer *r) { return !(r->last_opts.flags & RAST_FLAG_NO_SHADE); }

// Structure of a launch: whatever fixes its grids, block tables and kernel
// instances (the shading kernel has one per sampling mode).
static uint64_t plan_key(rast_renderer *const *rs, uint32_t n, int hidden)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
    mix(n);
    mix((uint64_t)hidden);
    mix((uint64_t)(n ? rs[0]->resolve_mode : 0));
    // which entries share a source mesh (vertex groups): only views share meshes
    bool views = false;
    for (uint32_t j = 0; j < n; ++j) views = views || rs[j]->scene_src != nullptr;
    std::unordered_map<const rast_mesh *, uint32_t> first;  // mesh -> first entry using it
    uint32_t entry = 0;
    for (uint32_t j = 0; j < n; ++j) {
        const rast_renderer *r = rs[j];
        mix((uint64_t)r->W << 32 | r->H);
        mix((uint64_t)r->last_opts.sample_state << 16 | (uint64_t)r->tile << 8 | (uint64_t)r->last_single << 4 |
            (uint64_t)job_learned(r) << 2 | job_shades(r));
        const std::vector<rast_mesh *> &ms = scene_meshes(const_cast<rast_renderer *>(r));
        mix(ms.size());
        for (const rast_mesh *m : ms) {
            mix((uint64_t)m->nv << 32 | m->ntri);
            if (views) mix(first.emplace(m, entry).first->second);
            ++entry;
        }
    }
    return h;
}

void plan_free(LaunchPlan &p)
{
    for (int i = 0; i < 2; ++i) {
        if (p.exec[i]) (void)hipGraphExecDestroy(p.exec[i]);
        p.exec[i] = nullptr;
    }
    for (int i = 0; i < 2; ++i) {
        if (p.side[i]) (void)hipStreamDestroy(p.side[i]);
        if (p.joined[i]) (void)hipEventDestroy(p.joined[i]);
        p.side[i] = nullptr;
        p.joined[i] = nullptr;
    }
    if (p.fork) (void)hipEventDestroy(p.fork);
    p.fork = nullptr;
    for (int i = 0; i < 2; ++i) {
        if (p.done[i]) (void)hipEventDestroy(p.done[i]);
        p.done[i] = nullptr;
        p.done_pending[i] = false;
        if (p.h_stage[i]) (void)hipHostFree(p.h_stage[i]);
        p.h_stage[i] = nullptr;
    }
    dfree(p.d_block_mesh);
    dfree(p.d_alive_part);
    dfree(p.d_tables);
    dfree(p.d_stage);
    dfree(p.d_sattr);
    p.built = false;
    p.key = 0;
}

// Waits for the last launch that used staging block k.
static int plan_wait_block(LaunchPlan &P, int k)
{
    if (P.done_pending[k]) {
        HIP_TRY(hipEventSynchronize(P.done[k]));
        P.done_pending[k] = false;
    }
    return RAST_OK;
}

// Every launch of the plan must be done before its tables change.
static int plan_wait(LaunchPlan &P)
{
    RAST_TRY(plan_wait_block(P, 0));
    return plan_wait_block(P, 1);
}

static int dev_alloc(DBuf &b, size_t bytes)
{
    if (b.p && b.n >= bytes) return RAST_OK;
    dfree(b);
    HIP_TRY(hipMalloc(&b.p, bytes < 256 ? 256 : bytes));
    b.n = bytes < 256 ? 256 : bytes;
    return RAST_OK;
}

// Grids, block -> job tables and block -> mesh tables of a launch structure.
static int build_plan(LaunchPlan &P, rast_renderer *const *rs, uint32_t n, int hidden, uint64_t key)
{
    RAST_TRY(plan_wait(P));
    for (int i = 0; i < 2; ++i) {
        if (P.exec[i]) (void)hipGraphExecDestroy(P.exec[i]);
        P.exec[i] = nullptr;
        // these events order work on the device and tell the host that a
        // launch is done with its staging block and tables; no host read of
        // results goes through them (counters and frames are read after a
        // stream synchronization), so they skip the system-scope fence and
        // its cache writeback (4K frame 1.4 us shorter)
        if (!P.done[i]) HIP_TRY(hipEventCreateWithFlags(&P.done[i], hipEventDisableTiming | hipEventDisableSystemFence));
    }
    P.built = false;
    if (!P.fork) {
        HIP_TRY(hipEventCreateWithFlags(&P.fork, hipEventDisableTiming | hipEventDisableSystemFence));
        for (int i = 0; i < 2; ++i) {
            HIP_TRY(hipStreamCreateWithFlags(&P.side[i], hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&P.joined[i], hipEventDisableTiming | hipEventDisableSystemFence));
        }
    }
    if (n > 65535) return rast_set_error(RAST_E_INVALID_ARG, "at most 65535 jobs per launch");
    Launch &L = P.L;
    L = Launch{};
    L.njobs = n;
    L.model_hidden = hidden;
    P.block0.assign((size_t)n * ST_COUNT, 0);
    P.nblocks.assign((size_t)n * ST_COUNT, 0);
    std::vector<uint16_t> tables[ST_COUNT];
    // vertex groups: the jobs' entries of one source mesh (views of one scene)
    // share vertex blocks, which read each vertex once for all of them
    std::vector<uint32_t> vs_group, tri_mesh, group_block0;
    std::vector<std::vector<uint32_t>> members;
    std::unordered_map<const rast_mesh *, uint32_t> group_of;
    P.vs_block0.clear();
    P.tri_block0.clear();
    uint64_t grid[ST_COUNT] = {};
    uint32_t mesh_entry = 0;
    for (uint32_t j = 0; j < n; ++j) {
        rast_renderer *r = rs[j];
        const uint32_t ntiles = ceil_div(r->W, r->tile) * ceil_div(r->H, r->tile);
        const uint32_t ntris = std::max<uint32_t>(r->total_tris, 1);
        const bool single = r->last_single;
        uint32_t nb[ST_COUNT] = {};
        nb[ST_INIT] = ceil_div(ntiles, 256);
        nb[ST_CAP] = single ? ceil_div(ntiles, kCapacityTilesPerBlock) : 0;
        nb[ST_FILL] = single ? 0 : std::min<uint32_t>(ceil_div(ntris, 256), 1024);
        // entries beyond capacity: placed by two passes for the previous-frame
        // policy, gathered by the resolve pass for learned capacities
        const bool placed = single && !job_learned(r);
        nb[ST_FBALLOC] = placed ? ceil_div(ntiles, 256) : 0;
        nb[ST_FBFILL] = placed ? std::min<uint32_t>(ceil_div(ntris, 256), 256) : 0;
        const int res = r->tile == 8 ? ST_RES8 : (r->tile == 16 ? ST_RES16 : ST_RES32);
        nb[res] = ntiles;
        nb[ST_SHADE] = job_shades(r) ? ceil_div(r->W, 64) * ceil_div(r->H, (uint32_t)kShadeRows) : 0;
        if (!nb[ST_SHADE]) ++L.unshaded_jobs;
        for (int st = 0; st < ST_COUNT; ++st) {
            P.block0[(size_t)j * ST_COUNT + st] = (uint32_t)grid[st];
            P.nblocks[(size_t)j * ST_COUNT + st] = nb[st];
            grid[st] += nb[st];
            if (n > 1) tables[st].insert(tables[st].end(), nb[st], (uint16_t)j);
        }
        if (single) ++L.cap_jobs;
        else ++L.exact_jobs;
        for (rast_mesh *m : scene_meshes(r)) {
            auto it = group_of.find(m);
            uint32_t g;
            if (it == group_of.end()) {
                g = (uint32_t)members.size();
                group_of.emplace(m, g);
                members.emplace_back();
                group_block0.push_back((uint32_t)vs_group.size());
                vs_group.insert(vs_group.end(), ceil_div(m->nv, kVertexChunk), g);
            } else {
                g = it->second;
            }
            members[g].push_back(mesh_entry);
            P.vs_block0.push_back(group_block0[g]);
            P.tri_block0.push_back((uint32_t)tri_mesh.size());
            tri_mesh.insert(tri_mesh.end(), ceil_div(m->ntri, 256), mesh_entry);
            ++mesh_entry;
        }
    }
    // a grid's threads must stay below 2^32 (256 threads per block in every stage but the scan)
    for (int st = 0; st < ST_COUNT; ++st)
        if (grid[st] >= (1ull << 24)) return rast_set_error(RAST_E_INVALID_ARG, "launch too large (stage %d)", st);
    if (vs_group.size() >= (1ull << 24) || tri_mesh.size() >= (1ull << 24))
        return rast_set_error(RAST_E_INVALID_ARG, "launch too large (meshes)");
    for (int st = 0; st < ST_COUNT; ++st) L.grid[st] = (uint32_t)grid[st];
    P.njobs = n;
    P.nmeshes = mesh_entry;
    L.vs_blocks = (uint32_t)vs_group.size();
    L.tri_blocks = (uint32_t)tri_mesh.size();
    // device tables: [vs block -> group][tri block -> mesh][groups (8-byte aligned)][group entries],
    // [stage tables]
    std::vector<uint32_t> tab(vs_group);
    tab.insert(tab.end(), tri_mesh.begin(), tri_mesh.end());
    if (tab.size() & 1) tab.push_back(0);
    const size_t groups_at = tab.size();
    tab.resize(groups_at + 2 * members.size());
    uint32_t ne = 0;
    for (size_t g = 0; g < members.size(); ++g) {
        tab[groups_at + 2 * g] = ne;
        tab[groups_at + 2 * g + 1] = (uint32_t)members[g].size();
        ne += (uint32_t)members[g].size();
    }
    const size_t entries_at = tab.size();
    for (const auto &mm : members) tab.insert(tab.end(), mm.begin(), mm.end());
    RAST_TRY(dev_alloc(P.d_block_mesh, (tab.size() + 2) * 4));
    RAST_TRY(dev_alloc(P.d_alive_part, ((size_t)L.tri_blocks + 1) * 4));
    uint32_t *bm = (uint32_t *)P.d_block_mesh.p;
    if (!tab.empty()) HIP_TRY(hipMemcpy(bm, tab.data(), tab.size() * 4, hipMemcpyHostToDevice));
    L.vs_block_group = bm;
    L.tri_block_mesh = bm + vs_group.size();
    L.vs_groups = reinterpret_cast<const uint2 *>(bm + groups_at);
    L.vs_entries = bm + entries_at;
    if (n > 1) {
        size_t total = 0;
        for (int st = 0; st < ST_COUNT; ++st) total += tables[st].size();
        RAST_TRY(dev_alloc(P.d_tables, (total + 8) * 2));
        uint16_t *tb = (uint16_t *)P.d_tables.p;
        size_t off = 0;
        for (int st = 0; st < ST_COUNT; ++st) {
            if (!tables[st].empty())
                HIP_TRY(hipMemcpy(tb + off, tables[st].data(), tables[st].size() * 2, hipMemcpyHostToDevice));
            L.table[st] = tables[st].empty() ? nullptr : tb + off;
            off += tables[st].size();
        }
    }
    // jobs drawing one scene (views of one owner): that scene's
    // view-independent attributes are written once, into the plan's buffer
    {
        std::unordered_map<const rast_renderer *, uint32_t> owner_jobs;
        for (uint32_t j = 0; j < n; ++j) ++owner_jobs[scene_owner(rs[j])];
        std::unordered_map<const rast_renderer *, size_t> owner_off;
        P.sattr_off.assign(n, SIZE_MAX);
        size_t sbytes = 0;
        for (uint32_t j = 0; j < n; ++j) {
            const rast_renderer *o = scene_owner(rs[j]);
            if (owner_jobs[o] < 2) continue;
            auto it = owner_off.find(o);
            if (it == owner_off.end()) {
                it = owner_off.emplace(o, sbytes).first;
                sbytes += ((size_t)rs[j]->total_verts + 1) * 32;
            }
            P.sattr_off[j] = it->second;
        }
        if (sbytes) RAST_TRY(dev_alloc(P.d_sattr, sbytes));
    }
    // shading instances: bit (sampling mode + 2 for the shared-scene layout)
    for (uint32_t j = 0; j < n; ++j)
        if (P.nblocks[(size_t)j * ST_COUNT + ST_SHADE])
            L.shade_modes |= 1u << (rs[j]->last_opts.sample_state + (P.sattr_off[j] != SIZE_MAX ? 2 : 0));
    // resolve form: the persistent kernels are sized to what the device holds
    L.resolve_mode = n ? rs[0]->resolve_mode : 0;
    L.fused_shade = 0;
    if (L.resolve_mode == RAST_RESOLVE_PERSISTENT_SHADE && (L.shade_modes == 1u || L.shade_modes == 2u))
        L.fused_shade = (int)L.shade_modes;
    for (int st : { ST_RES8, ST_RES16, ST_RES32 })
        L.resident[st] = L.resolve_mode && L.grid[st] ? (uint32_t)resolve_resident_blocks(st, L.fused_shade) : 0u;
    // staging block: the work counters (zeros), JobDesc[n], FrameCounters[n]
    // (zeros), then DevMesh[meshes]
    const size_t bytes = stage_meshes_offset(n) + (size_t)P.nmeshes * sizeof(DevMesh);
    if (bytes > P.stage_bytes || !P.h_stage[0] || !P.h_stage[1]) {
        for (int i = 0; i < 2; ++i) {
            if (P.h_stage[i]) (void)hipHostFree(P.h_stage[i]);
            P.h_stage[i] = nullptr;
            HIP_TRY(hipHostMalloc(&P.h_stage[i], bytes, hipHostMallocCoherent));  // read by k_upload
        }
    }
    RAST_TRY(dev_alloc(P.d_stage, bytes));
    P.stage_bytes = bytes;
    L.work = (uint32_t *)P.d_stage.p;
    L.jobs = (const JobDesc *)((char *)P.d_stage.p + kStageCounterBytes);
    L.meshes = (const DevMesh *)((char *)P.d_stage.p + stage_meshes_offset(n));
    P.key = key;
    P.built = true;
    return RAST_OK;
}

// Writes every job's descriptor and mesh entries (this frame's cameras,
// options and buffer parities) into the staging block of the next launch,
// once the launch before last, which used it, is done.
static int stage_launch(LaunchPlan &P, rast_renderer *const *rs, uint32_t n)
{
    RAST_TRY(plan_wait_block(P, P.parity));
    void *h = P.h_stage[P.parity];
    std::memset(h, 0, kStageCounterBytes);  // the launch's work counters start at zero
    // and so do the jobs' frame counters
    std::memset((char *)h + stage_counters_offset(n), 0, stage_meshes_offset(n) - stage_counters_offset(n));
    JobDesc *jd = (JobDesc *)((char *)h + kStageCounterBytes);
    DevMesh *dm = (DevMesh *)((char *)h + stage_meshes_offset(n));
    uint32_t me = 0;
    for (uint32_t j = 0; j < n; ++j) {
        rast_renderer *r = rs[j];
        const rast_frame_opts &o = r->last_opts;
        const std::vector<rast_mesh *> &meshes = scene_meshes(r);
        JobDesc &J = jd[j];
        memset(&J, 0, sizeof J);
        FrameParams &fp = J.fp;
        fp.width = r->W;
        fp.height = r->H;
        fp.cull_mode = o.cull_mode;
        fp.sample_state = o.sample_state;
        fp.visualisation = o.visualisation;
        fp.clear_rgba = o.clear_rgba;
        fp.tile = r->tile;
        fp.tiles_x = ceil_div(r->W, r->tile);
        fp.tiles_y = ceil_div(r->H, r->tile);
        fp.big_threshold = r->big_threshold;
        fp.micro_limit = r->micro_limit;
        fp.num_tris = r->total_tris;
        fp.num_verts = r->total_verts;
        fp.num_meshes = (uint32_t)meshes.size();
        memcpy(J.block0, &P.block0[(size_t)j * ST_COUNT], sizeof J.block0);
        memcpy(J.nblocks, &P.nblocks[(size_t)j * ST_COUNT], sizeof J.nblocks);
        J.mesh0 = me;
        J.single = r->last_single;
        J.learned = job_learned(r);
        J.small_limit = r->small_limit;
        J.margin = r->margin;
        J.slack = r->slack;
        J.queue_cap = r->queue_cap;
        J.pool_cap = r->pool_cap;
        J.big_cap = r->big_cap;
        J.ovf_cap = r->ovf_cap;
        // 1x1 stand-ins for missing maps (deviation D8); word 0 unused
        J.defaults[0] = 0xFFFFFFFFu;
        J.defaults[1] = o.default_normal;
        J.defaults[2] = o.default_specular;
        J.defaults[3] = o.default_gloss;
        J.model = J.learned ? rast_model_params(r->model) : nullptr;
        J.pos4 = (float4 *)r->pos4.p;
        J.attr = (float4 *)r->attr.p;
        J.trirec = (uint32_t *)r->trirec.p;
        J.alive = (uint8_t *)r->alive.p;
        J.tile_count = (uint32_t *)r->tile_count[r->cur].p;
        J.tile_off = (uint32_t *)r->tile_off.p;
        J.tile_cap = (uint32_t *)r->tile_cap.p;
        J.tile_pool = (uint32_t *)r->tile_pool.p;
        J.tile_cursor = (uint32_t *)r->tile_cursor.p;
        J.prev_count = (const uint32_t *)r->tile_count[r->cur ^ 1].p;
        J.prev_depth = (const float *)r->tile_min_depth[r->cur ^ 1].p;
        J.queue = (uint32_t *)r->queue.p;
        J.pool = (uint32_t *)r->pool.p;
        J.ovf = (uint2 *)r->ovf.p;
        J.survivors = (uint32_t *)r->survivors.p;
        J.big = (uint32_t *)r->big.p;
        J.big_plane = (float4 *)r->bigp.p;
        J.sattr = P.sattr_off[j] != SIZE_MAX ? (float4 *)((char *)P.d_sattr.p + P.sattr_off[j]) : nullptr;
        r->last_shared_attr = J.sattr != nullptr;
        J.counters = (FrameCounters *)((char *)P.d_stage.p + stage_counters_offset(n) + (size_t)j * sizeof(FrameCounters));
        J.host_counters = r->d_counters_host;
        J.alive_part = (uint32_t *)P.d_alive_part.p;
        J.alive_part0 = 0;
        J.alive_parts = 0;
        if (!meshes.empty()) {  // the job's setup blocks are contiguous, in mesh order
            J.alive_part0 = P.tri_block0[me];
            J.alive_parts = P.tri_block0[me + meshes.size() - 1] + ceil_div(meshes.back()->ntri, 256) - J.alive_part0;
        }
        J.tile_min_depth = (float *)r->tile_min_depth[r->cur].p;
        J.keys = (unsigned long long *)r->keys.p;
        J.vis = (uint32_t *)r->vis.p;
        J.depth = (float *)r->depth.p;
        J.rgba = (uint32_t *)r->rgba.p;
        const uint32_t *dflt = (const uint32_t *)((const char *)P.d_stage.p + kStageCounterBytes +
                                                  (size_t)j * sizeof(JobDesc) + offsetof(JobDesc, defaults));
        rast_frame_consts fc;
        camera_frame(r->last_cam, r->W, r->H, fc);
        uint32_t vb = 0;
        for (size_t i = 0; i < meshes.size(); ++i, ++me) {
            const rast_mesh *m = meshes[i];
            DevMesh &d = dm[me];
            memset(&d, 0, sizeof d);
            d.ivert = m->ivert14;
            d.idx = m->idx;
            d.num_verts = m->nv;
            d.num_tris = m->ntri;
            d.vert_base = vb;
            d.tri_base = (uint32_t)m->tri_base;
            d.vs_block0 = P.vs_block0[me];
            d.tri_block0 = P.tri_block0[me];
            d.topology = m->topology;
            d.has_diffuse = m->tex[0] != nullptr;
            d.job = j;
            // the rotated normal / tangent stays below attr_max times the
            // largest row sum of the world matrix's 3x3 part
            float rowsum = 0.f;
            for (int row = 0; row < 3; ++row)
                rowsum = std::max(rowsum, std::fabs(M4c(m->world, row, 0)) + std::fabs(M4c(m->world, row, 1)) +
                                              std::fabs(M4c(m->world, row, 2)));
            d.shade_exact = !(m->attr_max * rowsum <= 4.f);
            mat4_mul(fc.viewproj, m->world, d.wvp);  // CCR/CUDARenderer.cu:241
            memcpy(d.world, m->world, sizeof d.world);
            for (int q = 0; q < 3; ++q) {
                d.cam_pos[q] = fc.cam_pos[q];
                d.cam_fwd[q] = fc.cam_fwd[q];
            }
            for (int q = 0; q < 4; ++q) {
                if (m->tex[q]) d.tex[q] = { m->tex[q]->texels, m->tex[q]->w, m->tex[q]->h };
                else d.tex[q] = { dflt + q, 1u, 1u };
            }
            d.texq = m->texq;
            vb += m->nv;
        }
    }
    return RAST_OK;
}

// Timing events of one frame, recorded between the stages when RAST_FLAG_TIMING is set.
static int mark(rast_renderer *r, int which, hipStream_t s)
{
    if (!r) return RAST_OK;
    HIP_TRY(hipEventRecord(r->ev[which], s));
    r->ev_used[which] = 1;
    return RAST_OK;
}

// Everything a launch puts on a stream: the staging upload, one kernel per
// stage, the counters publish.  Graph-capturable (no memset, no host waits;
// events only when `timed` is set).
static int enqueue_launch(LaunchPlan &P, hipStream_t s, rast_renderer *timed)
{
    const Launch &L = P.L;
    HIP_TRY(launch_upload(P.d_stage.p, P.h_stage[P.parity], P.stage_bytes, s));
    RAST_TRY(mark(timed, EV_START, s));
    if (!timed) {
        // the capacities read only the previous frame's counts and depths:
        // their workgroups run in the vertex launch, beside the transform
        HIP_TRY(launch_vertex(L, 1, s));
    } else {  // one stage after the other (per-stage timing)
        HIP_TRY(launch_vertex(L, 0, s));
        RAST_TRY(mark(timed, EV_VERTEX, s));
        HIP_TRY(launch_vertex(L, 2, s));
        RAST_TRY(mark(timed, EV_MODEL, s));
    }
    HIP_TRY(launch_setup(L, s));
    RAST_TRY(mark(timed, EV_SETUP, s));
    HIP_TRY(launch_scan(L, s));
    RAST_TRY(mark(timed, EV_SCANCNT, s));
    HIP_TRY(launch_fill(L, s));
    RAST_TRY(mark(timed, EV_FILL, s));
    HIP_TRY(launch_fallback(L, s));
    RAST_TRY(mark(timed, EV_FALLBACK, s));
    {
        // tile-size groups are independent: the second and third run on side streams
        int groups[3], ng = 0;
        for (int st : { ST_RES16, ST_RES32, ST_RES8 })
            if (L.grid[st]) groups[ng++] = st;
        if (ng > 1) {
            HIP_TRY(hipEventRecord(P.fork, s));
            for (int i = 1; i < ng; ++i) HIP_TRY(hipStreamWaitEvent(P.side[i - 1], P.fork, 0));
        }
        for (int i = 0; i < ng; ++i) HIP_TRY(launch_resolve(L, groups[i], i == 0 ? s : P.side[i - 1]));
        for (int i = 1; i < ng; ++i) {
            HIP_TRY(hipEventRecord(P.joined[i - 1], P.side[i - 1]));
            HIP_TRY(hipStreamWaitEvent(s, P.joined[i - 1], 0));
        }
    }
    RAST_TRY(mark(timed, EV_RASTER, s));
    HIP_TRY(launch_shade(L, s));
    RAST_TRY(mark(timed, EV_SHADE, s));
    HIP_TRY(launch_publish(L, s));
    return RAST_OK;
}

static int capture_launch(LaunchPlan &P, hipStream_t s)
{
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_launch(P, s, nullptr);
    hipGraph_t cap = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &cap);
    if (rc != RAST_OK) {
        if (cap) (void)hipGraphDestroy(cap);
        return rc;
    }
    HIP_TRY(e);
    hipGraphExec_t &ex = P.exec[P.parity];
    const hipError_t ie = hipGraphInstantiate(&ex, cap, nullptr, nullptr, 0);
    (void)hipGraphDestroy(cap);
    if (ie != hipSuccess) ex = nullptr;
    HIP_TRY(ie);
    return RAST_OK;
}

// Runs a built and staged plan on s (replaying its graph in graph mode).
static int run_plan(LaunchPlan &P, hipStream_t s, bool graph, rast_renderer *timed)
{
    const int k = P.parity;
    if (graph) {
        if (!P.exec[k]) RAST_TRY(capture_launch(P, s));
        HIP_TRY(hipGraphLaunch(P.exec[k], s));
    } else {
        RAST_TRY(enqueue_launch(P, s, timed));
    }
    HIP_TRY(hipEventRecord(P.done[k], s));
    P.done_pending[k] = true;
    P.parity = k ^ 1;
    return RAST_OK;
}

static void finish_enqueue(rast_renderer *r)
{
    r->stats.num_triangles = r->total_tris;
    r->stats.num_tiles = ceil_div(r->W, r->tile) * ceil_div(r->H, r->tile);
    r->frame_pending = true;
    r->frame_valid = false;
    r->have_prev = true;
    r->cur ^= 1;
}

static int clear_keys_if_dirty(rast_renderer *r, hipStream_t s)
{
    if (!r->keys_dirty) return RAST_OK;
    HIP_TRY(launch_clear_keys((unsigned long long *)r->keys.p, (size_t)r->W * r->H, s));
    r->keys_dirty = false;
    return RAST_OK;
}

static int job_hidden(rast_renderer *const *rs, uint32_t n, int *out)
{
    int h = 0;
    for (uint32_t j = 0; j < n; ++j) {
        if (!job_learned(rs[j])) continue;
        const int hj = rast_model_hidden(rs[j]->model);
        if (h && hj != h) return rast_set_error(RAST_E_INVALID_ARG, "learned jobs of one launch must share the model width");
        h = hj;
    }
    *out = h;
    return RAST_OK;
}

static int enqueue_frame(rast_renderer *r)
{
    const rast_frame_opts &o = r->last_opts;
    RAST_TRY(prepare_job(r));
    const bool timing = (o.flags & RAST_FLAG_TIMING) != 0;
    const bool graph = (o.flags & RAST_FLAG_GRAPH) != 0 && !timing;
    r->timing = timing;
    memset(r->ev_used, 0, sizeof r->ev_used);
    int hidden = 0;
    RAST_TRY(job_hidden(&r, 1, &hidden));
    const uint64_t key = plan_key(&r, 1, hidden);
    if (!r->plan.built || r->plan.key != key) RAST_TRY(build_plan(r->plan, &r, 1, hidden, key));
    RAST_TRY(stage_launch(r->plan, &r, 1));
    r->stream_busy = true;
    RAST_TRY(clear_keys_if_dirty(r, r->stream));
    r->keys_dirty = true;  // until the frame (whose resolve pass resets the keys) is enqueued
    RAST_TRY(run_plan(r->plan, r->stream, graph, timing ? r : nullptr));
    r->keys_dirty = false;
    finish_enqueue(r);
    return RAST_OK;
}

// Largest queue storage (entries) that single-pass capacities alone may ask for.
constexpr uint32_t kQueueGrowLimit = 1u << 28;

// Re-run the last frame after growing whichever storage overflowed.
static int check_frame(rast_renderer *r)
{
    for (int attempt = 0; attempt < 4; ++attempt) {
        HIP_TRY(renderer_wait(r));
        const FrameCounters c = *r->h_counters;
        if (!c.storage_overflow) {
            r->frame_pending = false;
            r->frame_valid = true;
            // single-pass capacities beyond the queue storage were cut to fit
            // (the frame is exact all the same); the next frame gets room
            // for them, up to 2^28 entries
            if (r->last_single && c.capacity_wanted > r->queue_cap && r->queue_cap < kQueueGrowLimit)
                r->queue_cap = (uint32_t)std::min<uint64_t>((uint64_t)c.capacity_wanted * 5 / 4 + 1024, kQueueGrowLimit);
            return RAST_OK;
        }
        const uint64_t need_q = std::max<uint64_t>(c.total_pairs, c.capacity_total);
        if (need_q >= r->queue_cap) r->queue_cap = (uint32_t)std::min<uint64_t>(need_q + need_q / 4 + 1024, 0xFFFFFFF0ull);
        if (c.pool_used >= r->pool_cap) r->pool_cap = (uint32_t)std::min<uint64_t>((uint64_t)c.pool_used * 5 / 4 + 1024, 0xFFFFFFF0ull);
        if (c.num_big >= r->big_cap) r->big_cap = (uint32_t)std::min<uint64_t>((uint64_t)c.num_big * 5 / 4 + 1024, 0xFFFFFFF0ull);
        if (c.ovf_entries >= r->ovf_cap) r->ovf_cap = (uint32_t)std::min<uint64_t>((uint64_t)c.ovf_entries * 5 / 4 + 1024, 0x1FFFFFF0ull);
        r->cur ^= 1;  // redo the same frame into the same buffers
        r->have_prev = r->prev_valid_before;
        ++r->rerenders;
        RAST_TRY(enqueue_frame(r));
    }
    return rast_set_error(RAST_E_STATE, "frame storage kept overflowing");
}

extern "C" int rast_render(rast_renderer *r, const rast_camera *cam, const rast_frame_opts *opts)
{
    if (!r || !cam || !opts) return rast_set_error(RAST_E_INVALID_ARG, "rast_render: null argument");
    RAST_TRY(validate_opts(opts));
    RAST_TRY(validate_camera(cam));
    // A frame nobody read is not checked: the next one replaces it and takes
    // only capacity hints from it (its per-tile counts are exact even where
    // its queue storage overflowed), so the host does not wait for it.
    // Every frame that is read is checked first (rast_sync).
    r->last_cam = *cam;
    r->last_opts = *opts;
    r->prev_valid_before = r->have_prev;
    return enqueue_frame(r);
}

extern "C" int rast_sync(rast_renderer *r)
{
    if (!r) return rast_set_error(RAST_E_INVALID_ARG, "rast_sync: null renderer");
    HIP_TRY(hipSetDevice(r->device));
    if (r->frame_pending) return check_frame(r);
    HIP_TRY(renderer_wait(r));
    return RAST_OK;
}

extern "C" int rast_readback(rast_renderer *r, uint32_t *vis, float *depth, uint32_t *rgba)
{
    if (!r) return rast_set_error(RAST_E_INVALID_ARG, "rast_readback: null renderer");
    RAST_TRY(rast_sync(r));
    if (!r->frame_valid) return rast_set_error(RAST_E_STATE, "rast_readback: no rendered frame");
    const size_t bytes = (size_t)r->W * r->H * 4;
    if (vis) HIP_TRY(hipMemcpy(vis, r->vis.p, bytes, hipMemcpyDeviceToHost));
    if (depth) HIP_TRY(hipMemcpy(depth, r->depth.p, bytes, hipMemcpyDeviceToHost));
    if (rgba) HIP_TRY(hipMemcpy(rgba, r->rgba.p, bytes, hipMemcpyDeviceToHost));
    return RAST_OK;
}

static float elapsed(rast_renderer *r, int a, int b)
{
    if (!r->ev_used[a] || !r->ev_used[b]) return 0.f;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r->ev[a], r->ev[b]) != hipSuccess) return 0.f;
    return ms;
}

extern "C" int rast_get_stats(rast_renderer *r, rast_stats *out)
{
    if (!r || !out) return rast_set_error(RAST_E_INVALID_ARG, "rast_get_stats: null argument");
    RAST_TRY(rast_sync(r));
    const FrameCounters c = *r->h_counters;
    rast_stats s = r->stats;
    s.num_visible = c.num_alive;
    s.num_bin_pairs = c.total_pairs;
    s.num_big = c.num_big;
    s.overflow_tiles = c.overflow_tiles;
    s.overflow_entries = c.ovf_entries;
    s.capacity_total = r->last_single ? c.capacity_total : 0;
    s.rerenders = r->rerenders;
    s.ms_vertex = elapsed(r, EV_START, EV_VERTEX);
    s.ms_model = elapsed(r, EV_VERTEX, EV_MODEL);  // capacities and queue regions (single pass)
    s.ms_scan = elapsed(r, EV_SETUP, EV_SCANCNT);
    s.ms_setup = elapsed(r, EV_MODEL, EV_SETUP);
    s.ms_count = s.ms_setup;
    s.ms_fill = elapsed(r, EV_SCANCNT, EV_FILL);
    s.ms_fallback = elapsed(r, EV_FILL, EV_FALLBACK);  // only the previous-frame policy has these passes
    s.ms_raster = elapsed(r, EV_FALLBACK, EV_RASTER);
    s.ms_shade = elapsed(r, EV_RASTER, EV_SHADE);
    s.ms_total = elapsed(r, EV_START, EV_SHADE);
    *out = s;
    return RAST_OK;
}

extern "C" int rast_frame_buffers(rast_renderer *r, void **vis, void **depth, void **rgba)
{
    if (!r) return rast_set_error(RAST_E_INVALID_ARG, "rast_frame_buffers: null renderer");
    // the frame is finished and checked first (re-rendered if its queue
    // storage overflowed), as for rast_readback
    RAST_TRY(rast_sync(r));
    if (!r->frame_valid) return rast_set_error(RAST_E_STATE, "rast_frame_buffers: no rendered frame");
    if (vis) *vis = r->vis.p;
    if (depth) *depth = r->depth.p;
    if (rgba) *rgba = r->rgba.p;
    return RAST_OK;
}

extern "C" int rast_debug_read(rast_renderer *r, int which, void *dst, size_t bytes, size_t *size_out)
{
    if (!r) return rast_set_error(RAST_E_INVALID_ARG, "rast_debug_read: null renderer");
    RAST_TRY(rast_sync(r));
    const uint32_t ntiles = ceil_div(r->W, r->tile) * ceil_div(r->H, r->tile);
    const int last = r->cur ^ 1;  // buffers of the frame just rendered
    size_t size = 0;
    const void *src = nullptr;
    std::vector<uint32_t> tmp;
    switch (which) {
    case RAST_DEBUG_VERTEX_POS: size = (size_t)r->total_verts * 16; src = r->pos4.p; break;
    case RAST_DEBUG_VERTEX_ATTR:
        if (r->last_shared_attr)
            return rast_set_error(RAST_E_STATE, "vertex attributes of views rendered in one launch are split "
                                                "(view-independent part shared); render the view alone to read them");
        size = (size_t)r->total_verts * 48;
        src = r->attr.p;
        break;
    case RAST_DEBUG_TRI_RECORDS: size = (size_t)r->total_tris * kTriRecWords * 4; src = r->trirec.p; break;
    case RAST_DEBUG_TRI_ALIVE: size = r->total_tris; src = r->alive.p; break;
    case RAST_DEBUG_TILE_COUNTS: size = (size_t)ntiles * 4; src = r->tile_count[last].p; break;
    case RAST_DEBUG_TILE_CAPS: size = (size_t)ntiles * 4; src = r->tile_cap.p; break;
    case RAST_DEBUG_BIG_LIST: {
        const uint32_t nb = std::min(r->h_counters->num_big, r->big_cap);
        tmp.resize(nb + 1);
        tmp[0] = nb;
        if (nb) HIP_TRY(hipMemcpy(tmp.data() + 1, r->big.p, (size_t)nb * 4, hipMemcpyDeviceToHost));
        size = tmp.size() * 4;
        if (size_out) *size_out = size;
        if (dst) memcpy(dst, tmp.data(), std::min(size, bytes));
        return RAST_OK;
    }
    default: return rast_set_error(RAST_E_INVALID_ARG, "unknown debug buffer %d", which);
    }
    if (size_out) *size_out = size;
    if (dst && size && src) HIP_TRY(hipMemcpy(dst, src, std::min(size, bytes), hipMemcpyDeviceToHost));
    return RAST_OK;
}

extern "C" int rast_renderer_set_binmodel(rast_renderer *r, rast_binmodel *m, float margin, uint32_t slack)
{
    if (!r) return rast_set_error(RAST_E_INVALID_ARG, "rast_renderer_set_binmodel: null renderer");
    if (!(margin >= 1.f) || margin > 64.f) return rast_set_error(RAST_E_INVALID_ARG, "margin %g not in [1, 64]", margin);
    if (slack > (1u << 20)) return rast_set_error(RAST_E_INVALID_ARG, "slack %u above 2^20", slack);
    r->model = m;
    r->margin = margin;
    r->slack = slack;
    return RAST_OK;
}

// ------------------------------------------------------------ timer, hooks

extern "C" int rast_timer_start(rast_renderer *r)
{
    if (!r) return rast_set_error(RAST_E_INVALID_ARG, "rast_timer_start: null renderer");
    HIP_TRY(hipSetDevice(r->device));
    HIP_TRY(hipEventRecord(r->ev[EV_TIMER_START], r->stream));
    r->timer_started = true;
    return RAST_OK;
}

extern "C" int rast_timer_stop(rast_renderer *r, float *ms)
{
    if (!r || !ms) return rast_set_error(RAST_E_INVALID_ARG, "rast_timer_stop: null argument");
    if (!r->timer_started) return rast_set_error(RAST_E_STATE, "rast_timer_stop: timer not started");
    HIP_TRY(hipSetDevice(r->device));
    // a frame from a batch ran on the batch stream: the stop event follows it
    if (r->batch_seq) HIP_TRY(hipStreamWaitEvent(r->stream, g_bsync.done, 0));
    HIP_TRY(hipEventRecord(r->ev[EV_TIMER_STOP], r->stream));
    HIP_TRY(hipEventSynchronize(r->ev[EV_TIMER_STOP]));
    HIP_TRY(hipEventElapsedTime(ms, r->ev[EV_TIMER_START], r->ev[EV_TIMER_STOP]));
    return RAST_OK;
}

extern "C" int rast_draw_texture(rast_renderer *r, const rast_texture *t, int mode, int stretch, int sample_state)
{
    if (!r || !t) return rast_set_error(RAST_E_INVALID_ARG, "rast_draw_texture: null argument");
    if (!t->r || !t->texels) return rast_set_error(RAST_E_STATE, "rast_draw_texture: the texture's renderer was destroyed");
    if (mode != RAST_DRAW_TEXTURE && mode != RAST_DRAW_TEXTURE_GLOBAL)
        return rast_set_error(RAST_E_INVALID_ARG, "rast_draw_texture: unknown mode %d", mode);
    if (sample_state != RAST_SAMPLE_POINT && sample_state != RAST_SAMPLE_LINEAR)
        return rast_set_error(RAST_E_INVALID_ARG, "rast_draw_texture: sample_state %d", sample_state);
    HIP_TRY(hipSetDevice(r->device));
    if (r->frame_pending) RAST_TRY(check_frame(r));
    HIP_TRY(renderer_wait(r));  // earlier frames (also from batches) are done with the buffers
    HIP_TRY(launch_draw_texture((uint32_t *)r->vis.p, (float *)r->depth.p, (uint32_t *)r->rgba.p, r->W, r->H,
                                t->texels, t->w, t->h, mode, stretch != 0, sample_state, r->stream));
    r->stream_busy = true;
    r->frame_pending = false;
    r->frame_valid = true;
    return RAST_OK;
}

// ----------------------------------------------------------------- batching

// Launch plans of recent batch shapes (a few, least recently used replaced),
// and the stream the batches run on.
struct BatchCache {
    LaunchPlan plans[4];
    uint64_t last_use[4] = {};
    uint64_t clock = 0;
    hipStream_t stream = nullptr;
    int device = -1;
};
static BatchCache g_batch;
static std::mutex g_batch_mutex;

extern "C" int rast_render_batch(const rast_job *jobs, uint32_t n)
{
    if (!jobs || n == 0) return rast_set_error(RAST_E_INVALID_ARG, "rast_render_batch: no jobs");
    if (n > 65535) return rast_set_error(RAST_E_INVALID_ARG, "rast_render_batch: at most 65535 jobs");
    const int device = jobs[0].target ? jobs[0].target->device : 0;
    for (uint32_t i = 0; i < n; ++i) {
        rast_renderer *t = jobs[i].target;
        if (!t) return rast_set_error(RAST_E_INVALID_ARG, "job %u has no target renderer", i);
        if (t->device != device) return rast_set_error(RAST_E_INVALID_ARG, "job %u is on another device", i);
        RAST_TRY(validate_opts(&jobs[i].opts));
        RAST_TRY(validate_camera(&jobs[i].camera));
    }
    std::vector<rast_renderer *> rs(n);
    for (uint32_t i = 0; i < n; ++i) rs[i] = jobs[i].target;
    {
        std::vector<rast_renderer *> sorted(rs);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return rast_set_error(RAST_E_INVALID_ARG, "two jobs of a batch share a renderer");
    }
    std::lock_guard<std::mutex> lock(g_batch_mutex);
    HIP_TRY(hipSetDevice(device));
    if (g_batch.device != device) {
        if (g_batch.stream) {
            (void)hipStreamSynchronize(g_batch.stream);
            g_bsync.waited.store(g_bsync.recorded.load());
            for (LaunchPlan &p : g_batch.plans) plan_free(p);
            (void)hipStreamDestroy(g_batch.stream);
            (void)hipEventDestroy(g_bsync.done);
            g_batch.stream = nullptr;
            g_bsync.done = nullptr;
        }
        HIP_TRY(hipStreamCreateWithFlags(&g_batch.stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&g_bsync.done, hipEventDisableTiming));
        g_batch.device = device;
    }
    // host part of every job; each renderer's previous frame is checked first
    for (uint32_t i = 0; i < n; ++i) {
        rast_renderer *t = rs[i];
        t->last_cam = jobs[i].camera;  // an unread earlier frame is replaced unchecked, as in rast_render
        t->last_opts = jobs[i].opts;
        t->last_opts.flags &= ~RAST_FLAG_TIMING;
        t->prev_valid_before = t->have_prev;
        RAST_TRY(prepare_job(t));
        memset(t->ev_used, 0, sizeof t->ev_used);
    }
    int hidden = 0;
    RAST_TRY(job_hidden(rs.data(), n, &hidden));
    const uint64_t key = plan_key(rs.data(), n, hidden);
    int pick = -1, victim = 0;
    for (int k = 0; k < 4; ++k) {
        if (g_batch.plans[k].built && g_batch.plans[k].key == key) pick = k;
        if (g_batch.last_use[k] < g_batch.last_use[victim]) victim = k;
    }
    if (pick < 0) {
        pick = victim;
        RAST_TRY(build_plan(g_batch.plans[pick], rs.data(), n, hidden, key));
    }
    LaunchPlan &P = g_batch.plans[pick];
    g_batch.last_use[pick] = ++g_batch.clock;
    RAST_TRY(stage_launch(P, rs.data(), n));
    for (rast_renderer *t : rs) {
        if (t->stream_busy) {  // the batch follows earlier work on t
            HIP_TRY(hipEventRecord(t->join, t->stream));
            HIP_TRY(hipStreamWaitEvent(g_batch.stream, t->join, 0));
        }
        RAST_TRY(clear_keys_if_dirty(t, g_batch.stream));
        t->keys_dirty = true;  // until the batch (whose resolve pass resets the keys) is enqueued
    }
    RAST_TRY(run_plan(P, g_batch.stream, true, nullptr));
    const uint64_t seq = g_bsync.recorded.load() + 1;
    HIP_TRY(hipEventRecord(g_bsync.done, g_batch.stream));
    g_bsync.recorded.store(seq);
    for (rast_renderer *t : rs) {
        t->batch_seq = seq;  // later use of t waits for this batch (renderer_wait)
        t->keys_dirty = false;
        finish_enqueue(t);
    }
    return RAST_OK;
}
