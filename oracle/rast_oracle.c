/*
 * rast_oracle.c -- CPU restatement of the reference rasterizer path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline); the
 * product library never links or calls it.  See rast_oracle.h for the list of
 * reference files each part follows.  Build: -O2 -std=c99 -ffp-contract=off,
 * no fast-math, so each expression below is evaluated as written, one IEEE
 * binary32 operation at a time, left to right.
 *
 * Parity-contract deviations from the reference (all deterministic):
 *  D1 vis buffer = global triangle ID (mesh base + index-order triangle index),
 *     cleared to 0xFFFFFFFF.
 *  D2 depth fp32, cleared to 1.0; a fragment is accepted iff 0 <= z <= 1
 *     (NaN rejected) and !(z > depth); ties go to the later triangle; a
 *     stored zero is +0.0.
 *  D3 triangles are drawn in ascending global ID.
 *  D4 bbox = trunc(min)-1 .. trunc(max)+1 of the screen vertices with the
 *     float first clamped to [-2, W+2] (H for y), then clamped to [0,W]x[0,H].
 *  D5 cull order: back/front face (CCR/ERenderer.cpp:82-101, FrontFace uses
 *     >= 0), frustum reject (CCR/CUDAROPs.cu:147-178), then zero or
 *     non-finite screen-space area (precedent CCR/CUDARenderer.cu:1492-1496).
 *  D6 no truncating grids: every pixel is covered.
 *  D7 world matrix is used as given (no doubled translation).
 *  D8 shading follows the GPU formula (CCR/CUDAROPs.cu:472-514) with 1x1
 *     stand-ins for missing normal/specular/gloss maps; packing saturates
 *     (NaN -> 0), memory order R,G,B,A=255.
 *  D9 the camera projection keeps Identity()'s [3][3] = 1 exactly as
 *     CCR/Camera.cpp:97-108 leaves it (w = 1 - z_view): reproduced, not fixed.
 */
#include "rast_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI_F 3.14159265358979323846f /* (float)PI, CCR/MathUtilities.h:19 */

/* ---------------------------------------------------------------- helpers */

/* Reference matrices store data[col][row]; M(r,c) = data[c][r]
 * (CCR/Matrix.h:9-44).  Here: m[c*4 + r]. */
#define M4(m, r, c) ((m)[(c) * 4 + (r)])

/* CCR/Matrix4.h:120-148: out(r,c) = sum_k lm(r,k) * rm(k,c), k = 0..3 in order. */
void orc_mat4_mul(const float *lm, const float *rm, float *out)
{
    float t[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
            M4(t, r, c) = M4(lm, r, 0) * M4(rm, 0, c) + M4(lm, r, 1) * M4(rm, 1, c) +
                          M4(lm, r, 2) * M4(rm, 2, c) + M4(lm, r, 3) * M4(rm, 3, c);
    memcpy(out, t, sizeof t);
}

static void cross3(const float *a, const float *b, float *o)
{ /* CCR/Vector3.h:154-161 */
    float x = a[1] * b[2] - a[2] * b[1];
    float y = a[2] * b[0] - a[0] * b[2];
    float z = a[0] * b[1] - a[1] * b[0];
    o[0] = x; o[1] = y; o[2] = z;
}

static float dot3(const float *a, const float *b)
{ /* CCR/Vector3.h:150-152 */
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static void normalize3(float *v)
{ /* CCR/Vector.h:41-55: m = sqrt(Dot(v,v)); v *= 1/m; zero stays zero */
    float m = sqrtf(dot3(v, v));
    if (m == 0.f) { v[0] = v[1] = v[2] = 0.f; return; }
    float inv = 1.f / m;
    v[0] *= inv; v[1] *= inv; v[2] *= inv;
}

/* CCR/Matrix4.h:327-363 (FGED1 inverse), same operation order. */
void orc_mat4_inverse(const float *m, float *out)
{
    float a[3] = { m[0], m[1], m[2] };    /* column 0 .xyz */
    float b[3] = { m[4], m[5], m[6] };
    float c[3] = { m[8], m[9], m[10] };
    float d[3] = { m[12], m[13], m[14] };
    float x = M4(m, 3, 0), y = M4(m, 3, 1), z = M4(m, 3, 2), w = M4(m, 3, 3);
    float s[3], t[3], u[3], v[3];
    cross3(a, b, s);
    cross3(c, d, t);
    for (int i = 0; i < 3; ++i) {
        u[i] = a[i] * y - b[i] * x;
        v[i] = c[i] * w - d[i] * z;
    }
    float det = dot3(s, v) + dot3(t, u);
    float invDet = 1.0f / det;
    for (int i = 0; i < 3; ++i) { s[i] *= invDet; t[i] *= invDet; u[i] *= invDet; v[i] *= invDet; }
    float r0[3], r1[3], r2[3], r3[3], tmp[3];
    cross3(b, v, tmp); for (int i = 0; i < 3; ++i) r0[i] = tmp[i] + t[i] * y;
    cross3(v, a, tmp); for (int i = 0; i < 3; ++i) r1[i] = tmp[i] - t[i] * x;
    cross3(d, u, tmp); for (int i = 0; i < 3; ++i) r2[i] = tmp[i] + s[i] * w;
    cross3(u, c, tmp); for (int i = 0; i < 3; ++i) r3[i] = tmp[i] - s[i] * z;
    /* Matrix<4,4>(row-wise 16 values) constructor, CCR/Matrix4.h:22-31 */
    float rows[4][4] = {
        { r0[0], r0[1], r0[2], -dot3(b, t) },
        { r1[0], r1[1], r1[2], dot3(a, t) },
        { r2[0], r2[1], r2[2], -dot3(d, s) },
        { r3[0], r3[1], r3[2], dot3(c, s) },
    };
    for (int r = 0; r < 4; ++r)
        for (int cc = 0; cc < 4; ++cc) M4(out, r, cc) = rows[r][cc];
}

/* CCR/Camera.cpp:6-31,97-117 and CCR/CUDARenderer.cu:212-217. */
void orc_camera_frame(const orc_camera *cam, uint32_t width, uint32_t height, orc_frame *fr)
{
    static const float worldUp[3] = { 0.f, 1.f, 0.f };
    float fwd[3] = { cam->forward[0], cam->forward[1], cam->forward[2] };
    float right[3], up[3];
    /* RecalculateVectors */
    normalize3(fwd);
    cross3(worldUp, fwd, right);
    normalize3(right);
    cross3(fwd, right, up);
    normalize3(up);
    float lookAt[16];
    memset(lookAt, 0, sizeof lookAt);
    for (int i = 0; i < 3; ++i) {
        lookAt[0 + i] = right[i];
        lookAt[4 + i] = up[i];
        lookAt[8 + i] = fwd[i];
        lookAt[12 + i] = cam->position[i];
    }
    lookAt[15] = 1.f;
    /* SetFov: tanf(ToRadians(deg) / 2), ToRadians = deg * ((float)PI / 180.f) */
    float radians = cam->fov_deg * (ORC_PI_F / 180.f);
    float fov = tanf(radians / 2.f);
    float aspect = (float)width / (float)height;
    float n = cam->near_z, f = cam->far_z;
    float A = f / (n - f);
    float B = (f * n) / (n - f);
    float value = 1.f / (aspect * fov);
    float invFov = 1.f / fov;
    float *P = fr->proj;
    memset(P, 0, 16 * sizeof(float));
    P[0] = 1.f; P[5] = 1.f; P[10] = 1.f; P[15] = 1.f; /* Identity(); [3][3] stays 1 (D9) */
    P[0 * 4 + 0] = value;
    P[1 * 4 + 1] = invFov;
    P[2 * 4 + 2] = A;
    P[2 * 4 + 3] = -1.f;
    P[3 * 4 + 2] = B;
    orc_mat4_inverse(lookAt, fr->view);
    orc_mat4_mul(fr->proj, fr->view, fr->viewproj);
    for (int i = 0; i < 3; ++i) {
        fr->cam_pos[i] = cam->position[i];
        fr->cam_fwd[i] = fwd[i];
        fr->cam_right[i] = right[i];
        fr->cam_up[i] = up[i];
    }
}

/* ------------------------------------------------------- vertex transform */

/* CCR/CUDAROPs.cu:99-118 (GetNDCVertex), CCR/Matrix4.h:158-166 (M * Point4,
 * the point's w is not read: + m(r,3)), CCR/Matrix3.h:126-133 (R * v). */
static void transform_vertex(const orc_ivertex *iv, const float *wvp, const float *world,
                             const float *cam_pos, float *pos4, float *attr)
{
    const float *p = iv->p;
    float o[4];
    for (int r = 0; r < 4; ++r)
        o[r] = M4(wvp, r, 0) * p[0] + M4(wvp, r, 1) * p[1] + M4(wvp, r, 2) * p[2] + M4(wvp, r, 3);
    o[0] /= o[3];
    o[1] /= o[3];
    o[2] /= o[3];
    pos4[0] = o[0]; pos4[1] = o[1]; pos4[2] = o[2]; pos4[3] = o[3];
    /* rotation = (FMatrix3)world: top-left 3x3, same (r,c) indexing */
    float nn[3], tt[3], wp[3], vd[3];
    for (int r = 0; r < 3; ++r) {
        nn[r] = M4(world, r, 0) * iv->n[0] + M4(world, r, 1) * iv->n[1] + M4(world, r, 2) * iv->n[2];
        tt[r] = M4(world, r, 0) * iv->tan[0] + M4(world, r, 1) * iv->tan[1] + M4(world, r, 2) * iv->tan[2];
        wp[r] = M4(world, r, 0) * p[0] + M4(world, r, 1) * p[1] + M4(world, r, 2) * p[2] + M4(world, r, 3);
    }
    for (int i = 0; i < 3; ++i) vd[i] = wp[i] - cam_pos[i];
    normalize3(vd);
    attr[0] = iv->uv[0]; attr[1] = iv->uv[1];
    attr[2] = nn[0]; attr[3] = nn[1]; attr[4] = nn[2];
    attr[5] = tt[0]; attr[6] = tt[1]; attr[7] = tt[2];
    attr[8] = vd[0]; attr[9] = vd[1]; attr[10] = vd[2];
    attr[11] = iv->c[0]; attr[12] = iv->c[1]; attr[13] = iv->c[2];
    attr[14] = 1.f / o[3]; /* invDepth, CCR/CUDAROPs.cu:347-350 */
    attr[15] = 0.f;
}

void orc_vertex_transform(const orc_mesh *mesh, const orc_frame *fr, float *out_pos4, float *out_attr)
{
    float wvp[16];
    orc_mat4_mul(fr->viewproj, mesh->world, wvp); /* CCR/CUDARenderer.cu:241 */
    for (uint32_t i = 0; i < mesh->num_verts; ++i)
        transform_vertex(&mesh->verts[i], wvp, mesh->world, fr->cam_pos, out_pos4 + 4 * (size_t)i,
                         out_attr + 16 * (size_t)i);
}

/* ------------------------------------------------------------ triangles */

static uint32_t mesh_num_triangles(const orc_mesh *m)
{ /* CCR/CUDAMesh.cpp:27-40 */
    if (m->topology == ORC_TOPO_LIST) return m->num_indices / 3;
    return m->num_indices >= 3 ? m->num_indices - 2 : 0;
}

/* Triangle k of a mesh: list (CCR/ERenderer.cpp:75-79) or strip.  Strips
 * follow the GPU path being replaced, TA_Kernel (CCR/CUDARenderer.cu:733-746):
 * triangle k takes indices k, k+1, k+2 and odd k swap the last two.
 * Deviation D10: ERenderer's CPU loop (CCR/ERenderer.cpp:115-157) starts
 * with the swap (its first triangle uses k, k+2, k+1) and toggles it only
 * after a triangle passed culling, so a triangle's winding there depends on
 * how many earlier triangles were culled; the index-parity rule makes every
 * triangle independent (and is what the CUDA path intends: its strip branch
 * declares a second `triangle` that shadows the one it culls and stores). */
static void triangle_indices(const orc_mesh *m, uint32_t k, uint32_t idx[3])
{
    if (m->topology == ORC_TOPO_LIST) {
        idx[0] = m->indices[3 * (size_t)k];
        idx[1] = m->indices[3 * (size_t)k + 1];
        idx[2] = m->indices[3 * (size_t)k + 2];
    } else {
        int odd = (int)(k & 1u);
        idx[0] = m->indices[k];
        idx[1] = odd ? m->indices[k + 2] : m->indices[k + 1];
        idx[2] = odd ? m->indices[k + 1] : m->indices[k + 2];
    }
}

static int all_outside(float a, float b, float c, float lo, float hi)
{ /* CCR/ERenderer.cpp:556-572 */
    return (a < lo && b < lo && c < lo) || (a > hi && b > hi && c > hi);
}

static int clamp_trunc(float v, float hi)
{ /* D4: clamp in float, then truncate toward zero (the reference's (short) cast) */
    float c = v < -2.f ? -2.f : v;
    c = c > hi ? hi : c;
    return (int)c;
}

typedef struct {
    float sx[3], sy[3], z[3];
    float area;                 /* |Cross(v0 - v1, v0 - v2)| in screen space */
    int xmin, ymin, xmax, ymax; /* pixel range [xmin,xmax) x [ymin,ymax) */
} tri_setup;

/* Returns 1 if the triangle survives (D5), filling *ts. */
static int setup_triangle(const float *P0, const float *P1, const float *P2, const float *cam_fwd,
                          int cull_mode, uint32_t W, uint32_t H, tri_setup *ts)
{
    if (cull_mode != ORC_CULL_NONE) { /* CCR/ERenderer.cpp:82-101 */
        float e1[3] = { P1[0] - P0[0], P1[1] - P0[1], P1[2] - P0[2] };
        float e2[3] = { P2[0] - P0[0], P2[1] - P0[1], P2[2] - P0[2] };
        float fn[3];
        cross3(e1, e2, fn);
        normalize3(fn);
        float cv = dot3(cam_fwd, fn);
        if (cull_mode == ORC_CULL_BACK) { if (cv <= 0.f) return 0; }
        else { if (cv >= 0.f) return 0; }
    }
    /* IsTriangleVisible, CCR/ERenderer.cpp:574-584 */
    if (all_outside(P0[0], P1[0], P2[0], -1.f, 1.f)) return 0;
    if (all_outside(P0[1], P1[1], P2[1], -1.f, 1.f)) return 0;
    if (all_outside(P0[2], P1[2], P2[2], 0.f, 1.f)) return 0;
    /* NDCToScreenSpace, CCR/ERenderer.cpp:603-607 */
    const float *Pv[3] = { P0, P1, P2 };
    for (int i = 0; i < 3; ++i) {
        ts->sx[i] = ((Pv[i][0] + 1) / 2) * (float)W;
        ts->sy[i] = ((1 - Pv[i][1]) / 2) * (float)H;
        ts->z[i] = Pv[i][2];
    }
    /* area as in IsPixelInTriangle: |Cross(v0 - v1, v2 - v0)| (CCR/ERenderer.cpp:503-508) */
    float eAx = ts->sx[0] - ts->sx[1], eAy = ts->sy[0] - ts->sy[1];
    float eCx = ts->sx[2] - ts->sx[0], eCy = ts->sy[2] - ts->sy[0];
    float cr = eAx * eCy - eAy * eCx;
    if (!(cr != 0.f) || !isfinite(cr)) return 0;
    ts->area = fabsf(cr);
    /* GetBoundingBox, CCR/ERenderer.cpp:609-624 with D4 */
    float mnx = ts->sx[0], mxx = ts->sx[0], mny = ts->sy[0], mxy = ts->sy[0];
    for (int i = 1; i < 3; ++i) {
        if (ts->sx[i] < mnx) mnx = ts->sx[i];
        if (ts->sx[i] > mxx) mxx = ts->sx[i];
        if (ts->sy[i] < mny) mny = ts->sy[i];
        if (ts->sy[i] > mxy) mxy = ts->sy[i];
    }
    int xmin = clamp_trunc(mnx, (float)W + 2.f) - 1;
    int ymin = clamp_trunc(mny, (float)H + 2.f) - 1;
    int xmax = clamp_trunc(mxx, (float)W + 2.f) + 1;
    int ymax = clamp_trunc(mxy, (float)H + 2.f) + 1;
    if (xmin < 0) xmin = 0;
    if (ymin < 0) ymin = 0;
    if (xmax > (int)W) xmax = (int)W;
    if (ymax > (int)H) ymax = (int)H;
    ts->xmin = xmin; ts->ymin = ymin; ts->xmax = xmax; ts->ymax = ymax;
    return 1;
}

/* IsPixelInTriangle, CPU form (CCR/ERenderer.cpp:497-542). */
static int pixel_in_triangle(const tri_setup *t, float px, float py, float w[3])
{
    float eAx = t->sx[0] - t->sx[1], eAy = t->sy[0] - t->sy[1];
    float eBx = t->sx[1] - t->sx[2], eBy = t->sy[1] - t->sy[2];
    float eCx = t->sx[2] - t->sx[0], eCy = t->sy[2] - t->sy[0];
    float c;
    c = eAx * (py - t->sy[0]) - eAy * (px - t->sx[0]);
    if (c < 0.f) return 0;
    w[2] = c / t->area;
    c = eBx * (py - t->sy[1]) - eBy * (px - t->sx[1]);
    if (c < 0.f) return 0;
    w[0] = c / t->area;
    c = eCx * (py - t->sy[2]) - eCy * (px - t->sx[2]);
    if (c < 0.f) return 0;
    w[1] = c / t->area;
    return 1;
}

/* DepthTest's interpolation (CCR/ERenderer.cpp:546) and the D2 accept window. */
static int fragment_depth(const tri_setup *t, const float w[3], float *z)
{
    float zi = (w[0] * t->z[0]) + (w[1] * t->z[1]) + (w[2] * t->z[2]);
    if (!(zi >= 0.f && zi <= 1.f)) return 0;
    *z = zi == 0.f ? 0.f : zi; /* canonical +0.0 */
    return 1;
}

/* ---------------------------------------------------------------- scene */

typedef struct {
    const orc_mesh *meshes;
    int num_meshes;
    uint64_t *tri_base;      /* [num_meshes + 1] global triangle ID bases */
    float **pos4, **attr;    /* per mesh VS outputs */
    uint8_t *alive;          /* [num_triangles] */
    tri_setup *ts;           /* [num_triangles] */
    uint64_t num_tris;
    orc_frame fr;
    uint32_t W, H;
    int cull_mode;
} scene_state;

static void scene_free(scene_state *s)
{
    if (s->pos4) for (int m = 0; m < s->num_meshes; ++m) free(s->pos4[m]);
    if (s->attr) for (int m = 0; m < s->num_meshes; ++m) free(s->attr[m]);
    free(s->pos4); free(s->attr); free(s->tri_base); free(s->alive); free(s->ts);
    memset(s, 0, sizeof *s);
}

typedef struct { scene_state *s; uint64_t t0, t1; } setup_job;

static int locate_mesh(const scene_state *s, uint64_t gid, uint32_t *local)
{
    int lo = 0, hi = s->num_meshes - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (s->tri_base[mid] <= gid) lo = mid; else hi = mid - 1;
    }
    *local = (uint32_t)(gid - s->tri_base[lo]);
    return lo;
}

static void *setup_worker(void *arg)
{
    setup_job *j = (setup_job *)arg;
    scene_state *s = j->s;
    for (uint64_t g = j->t0; g < j->t1; ++g) {
        uint32_t k;
        int m = locate_mesh(s, g, &k);
        const orc_mesh *mesh = &s->meshes[m];
        uint32_t idx[3];
        triangle_indices(mesh, k, idx);
        const float *P0 = s->pos4[m] + 4 * (size_t)idx[0];
        const float *P1 = s->pos4[m] + 4 * (size_t)idx[1];
        const float *P2 = s->pos4[m] + 4 * (size_t)idx[2];
        s->alive[g] = (uint8_t)setup_triangle(P0, P1, P2, s->fr.cam_fwd, s->cull_mode, s->W, s->H, &s->ts[g]);
    }
    return NULL;
}

static int run_parallel(void *(*fn)(void *), void *jobs, size_t job_size, int n)
{
    if (n <= 1) { fn(jobs); return 0; }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n);
    if (!th) return -1;
    int started = 0;
    for (int i = 0; i < n; ++i) {
        if (pthread_create(&th[i], NULL, fn, (char *)jobs + job_size * (size_t)i) != 0) {
            fn((char *)jobs + job_size * (size_t)i);
            th[i] = 0;
        } else {
            ++started;
        }
    }
    for (int i = 0; i < n; ++i) if (th[i]) pthread_join(th[i], NULL);
    free(th);
    (void)started;
    return 0;
}

static int scene_build(scene_state *s, const orc_mesh *meshes, int num_meshes, const orc_camera *cam,
                       int cull_mode, uint32_t W, uint32_t H, int nthreads)
{
    memset(s, 0, sizeof *s);
    s->meshes = meshes; s->num_meshes = num_meshes; s->W = W; s->H = H; s->cull_mode = cull_mode;
    orc_camera_frame(cam, W, H, &s->fr);
    s->tri_base = (uint64_t *)calloc((size_t)num_meshes + 1, sizeof(uint64_t));
    s->pos4 = (float **)calloc((size_t)num_meshes, sizeof(float *));
    s->attr = (float **)calloc((size_t)num_meshes, sizeof(float *));
    if (!s->tri_base || !s->pos4 || !s->attr) { scene_free(s); return -1; }
    for (int m = 0; m < num_meshes; ++m) {
        s->tri_base[m + 1] = s->tri_base[m] + mesh_num_triangles(&meshes[m]);
        s->pos4[m] = (float *)malloc(sizeof(float) * 4 * ((size_t)meshes[m].num_verts + 1));
        s->attr[m] = (float *)malloc(sizeof(float) * 16 * ((size_t)meshes[m].num_verts + 1));
        if (!s->pos4[m] || !s->attr[m]) { scene_free(s); return -1; }
        orc_vertex_transform(&meshes[m], &s->fr, s->pos4[m], s->attr[m]);
    }
    s->num_tris = s->tri_base[num_meshes];
    s->alive = (uint8_t *)malloc(s->num_tris + 1);
    s->ts = (tri_setup *)malloc(sizeof(tri_setup) * (s->num_tris + 1));
    if (!s->alive || !s->ts) { scene_free(s); return -1; }
    if (nthreads < 1) nthreads = 1;
    setup_job *jobs = (setup_job *)malloc(sizeof(setup_job) * (size_t)nthreads);
    if (!jobs) { scene_free(s); return -1; }
    for (int i = 0; i < nthreads; ++i) {
        jobs[i].s = s;
        jobs[i].t0 = s->num_tris * (uint64_t)i / (uint64_t)nthreads;
        jobs[i].t1 = s->num_tris * (uint64_t)(i + 1) / (uint64_t)nthreads;
    }
    run_parallel(setup_worker, jobs, sizeof(setup_job), nthreads);
    free(jobs);
    return 0;
}

int orc_setup(const orc_mesh *meshes, int num_meshes, const orc_frame *fr, int cull_mode,
              uint32_t W, uint32_t H, uint8_t *out_flags, uint32_t *out_rec)
{
    uint64_t g = 0;
    for (int m = 0; m < num_meshes; ++m) {
        const orc_mesh *mesh = &meshes[m];
        float *pos4 = (float *)malloc(sizeof(float) * 4 * ((size_t)mesh->num_verts + 1));
        float *attr = (float *)malloc(sizeof(float) * 16 * ((size_t)mesh->num_verts + 1));
        if (!pos4 || !attr) { free(pos4); free(attr); return -1; }
        orc_vertex_transform(mesh, fr, pos4, attr);
        uint32_t nt = mesh_num_triangles(mesh);
        for (uint32_t k = 0; k < nt; ++k, ++g) {
            uint32_t idx[3];
            triangle_indices(mesh, k, idx);
            tri_setup ts;
            int ok = setup_triangle(pos4 + 4 * (size_t)idx[0], pos4 + 4 * (size_t)idx[1],
                                    pos4 + 4 * (size_t)idx[2], fr->cam_fwd, cull_mode, W, H, &ts);
            out_flags[g] = (uint8_t)ok;
            uint32_t *r = out_rec + 12 * g;
            memset(r, 0, 12 * sizeof(uint32_t));
            if (ok) {
                float f[10] = { ts.sx[0], ts.sy[0], ts.sx[1], ts.sy[1], ts.sx[2], ts.sy[2],
                                ts.z[0], ts.z[1], ts.z[2], ts.area };
                memcpy(r, f, sizeof f);
                r[10] = (uint32_t)ts.xmin | ((uint32_t)ts.ymin << 16);
                r[11] = (uint32_t)ts.xmax | ((uint32_t)ts.ymax << 16);
            }
        }
        free(pos4); free(attr);
    }
    return 0;
}

/* ------------------------------------------------------------- raster */

typedef struct { scene_state *s; uint32_t y0, y1; uint32_t *vis; float *depth; } raster_job;

/* Sequential ERenderer loop (CCR/ERenderer.cpp:59-158, 269-295, 544-554)
 * restricted to the rows [y0, y1): per triangle in ascending global ID, per
 * pixel in its bbox row-major, the sample point is the integer pixel. */
static void *raster_worker(void *arg)
{
    raster_job *j = (raster_job *)arg;
    scene_state *s = j->s;
    const uint32_t W = s->W;
    for (uint64_t g = 0; g < s->num_tris; ++g) {
        if (!s->alive[g]) continue;
        const tri_setup *t = &s->ts[g];
        int y0 = t->ymin > (int)j->y0 ? t->ymin : (int)j->y0;
        int y1 = t->ymax < (int)j->y1 ? t->ymax : (int)j->y1;
        for (int r = y0; r < y1; ++r) {
            for (int c = t->xmin; c < t->xmax; ++c) {
                float w[3], z;
                if (!pixel_in_triangle(t, (float)c, (float)r, w)) continue;
                if (!fragment_depth(t, w, &z)) continue;
                size_t pid = (size_t)c + (size_t)r * W;
                if (z > j->depth[pid]) continue;
                j->depth[pid] = z;
                j->vis[pid] = (uint32_t)g;
            }
        }
    }
    return NULL;
}

/* ------------------------------------------------------------ shading */

static void texel_rgb(uint32_t t, float rgb[3])
{ /* RGBA::GetRGBColor, CCR/RGBRaw.h:34-38 */
    rgb[0] = (float)(t & 0xFFu) / 255.f;
    rgb[1] = (float)((t >> 8) & 0xFFu) / 255.f;
    rgb[2] = (float)((t >> 16) & 0xFFu) / 255.f;
}

/* int(uv * w + 0.5f) with the reference's out-of-range test, defined for all
 * inputs: valid iff -1 < f < w (then the truncation lands in [0, w)). */
static int texel_coord(float uv, uint32_t n, int *out)
{
    float f = uv * (float)n + 0.5f;
    if (!(f > -1.f && f < (float)n)) return 0;
    *out = (int)f;
    return 1;
}

static void sample_record(const orc_texture *t, const float uv[2], int mode, float rgb[3],
                          uint8_t *bitmap)
{
    int x, y;
    if (!texel_coord(uv[0], t->w, &x) || !texel_coord(uv[1], t->h, &y)) {
        rgb[0] = 1.f; rgb[1] = 0.f; rgb[2] = 1.f; /* magenta, CCR/CUDATextureSampler.cuh:97-98 */
        return;
    }
    int w = (int)t->w;
    int orig = x + y * w;
    if (mode == ORC_SAMPLE_POINT) { /* CCR/CUDATextureSampler.cuh:86-104 */
        if (bitmap) bitmap[orig] = 1;
        texel_rgb(t->texels[orig], rgb);
        return;
    }
    /* "Linear": CCR/CUDATextureSampler.cuh:112-190 */
    int texSize = w * (int)t->h;
    int nb[4];
    nb[0] = orig - 1; if (nb[0] < 0) nb[0] = orig;
    nb[1] = orig + 1; if (nb[1] >= texSize) nb[1] = orig;
    nb[2] = orig + w; if (nb[2] >= texSize) nb[2] = orig;
    nb[3] = orig - w; if (nb[3] < 0) nb[3] = orig;
    const float weight = 1.f / (4 * 2);
    float acc[3] = { 0.f, 0.f, 0.f };
    for (int i = 0; i < 4; ++i) {
        uint32_t tx = t->texels[nb[i]];
        if (bitmap) bitmap[nb[i]] = 1;
        acc[0] += (float)(tx & 0xFFu) * weight;
        acc[1] += (float)((tx >> 8) & 0xFFu) * weight;
        acc[2] += (float)((tx >> 16) & 0xFFu) * weight;
    }
    uint32_t to = t->texels[orig];
    if (bitmap) bitmap[orig] = 1;
    acc[0] += (float)(to & 0xFFu) / 2.f;
    acc[1] += (float)((to >> 8) & 0xFFu) / 2.f;
    acc[2] += (float)((to >> 16) & 0xFFu) / 2.f;
    const float rev = 1.0f / 255.f; /* RGBColor::operator/(float), CCR/RGBColor.h:38-42 */
    rgb[0] = acc[0] * rev; rgb[1] = acc[1] * rev; rgb[2] = acc[2] * rev;
}

void orc_sample(const orc_texture *t, const float uv[2], int sample_state, float rgb[3])
{
    sample_record(t, uv, sample_state, rgb, NULL);
}

static uint8_t pack_channel(float c)
{ /* (unsigned char)(c * 255) with CUDA's saturating float->u8 conversion */
    float v = c * 255.f;
    if (!(v > 0.f)) return 0;
    if (v >= 255.f) return 255;
    return (uint8_t)v;
}

uint32_t orc_pack_rgba(const float rgb[3])
{
    return (uint32_t)pack_channel(rgb[0]) | ((uint32_t)pack_channel(rgb[1]) << 8) |
           ((uint32_t)pack_channel(rgb[2]) << 16) | 0xFF000000u;
}

/* ShadePixelUnsafe, CCR/CUDAROPs.cu:472-514 (isFlipGreenChannel = false). */
static void shade_pbr(const float diff[3], const float nrm[3], const float spc[3], const float gls[3],
                      const float n[3], const float tan[3], const float vd[3], float out[3])
{
    static const float ambient = 0.05f;
    static const float L[3] = { 0.577f, -0.577f, -0.577f };
    const float lightIntensity = 7.0f, shininess = 25.f;
    float cr[3];
    cross3(tan, n, cr);
    float bn[3] = { -cr[0], -cr[1], -cr[2] };
    float fnl[3] = { 2.f * nrm[0] - 1.f, 2.f * nrm[1] - 1.f, 2.f * nrm[2] - 1.f };
    float N[3];
    for (int r = 0; r < 3; ++r) N[r] = tan[r] * fnl[0] + bn[r] * fnl[1] + n[r] * fnl[2];
    float negN[3] = { -N[0], -N[1], -N[2] };
    float oa = dot3(negN, L);
    /* ClampFast(observedArea, 0.f, observedArea) returns its argument unchanged */
    oa /= ORC_PI_F;
    oa *= lightIntensity;
    float dcol[3] = { diff[0] * oa, diff[1] * oa, diff[2] * oa };
    /* Reflect(L, N) = -L + (2 * Dot(L, N)) * N, CCR/Vector.h:76-78 */
    float twoDot = 2.f * dot3(L, N);
    float refl[3];
    for (int i = 0; i < 3; ++i) refl[i] = -L[i] + twoDot * N[i];
    float angle = dot3(vd, refl);
    angle = angle < 0.f ? 0.f : angle;
    angle = angle > 1.f ? 1.f : angle;
    angle = powf(angle, gls[0] * shininess);
    float scol[3] = { spc[0] * angle, spc[1] * angle, spc[2] * angle };
    for (int i = 0; i < 3; ++i) out[i] = ambient + dcol[i] + scol[i];
    /* MaxToOne, CCR/RGBColor.h:74-79 with Max(a,b) = a < b ? b : a */
    float gb = out[1] < out[2] ? out[2] : out[1];
    float mx = out[0] < gb ? gb : out[0];
    if (mx > 1.f) {
        float rev = 1.0f / mx;
        out[0] *= rev; out[1] *= rev; out[2] *= rev;
    }
}

typedef struct {
    scene_state *s;
    const orc_opts *opts;
    uint32_t y0, y1;
    const uint32_t *vis;
    const float *depth;
    uint32_t *rgba;
    uint8_t **tex_bitmap;  /* per mesh per slot -> bitmap (shared between meshes using one texture) */
    uint8_t **vert_bitmap; /* per mesh */
    const orc_texture *defaults;
} shade_job;

/* Interpolation of RasterizePixelAtomic (CCR/CUDAROPs.cu:329-380) at pixel
 * (x, y) of triangle g, whose vertices have the VS outputs of s; the weights
 * use the GPU edge form (CCR/CUDAROPs.cu:120-145), which is bitwise the CPU
 * form.  uv and n are perspective-correct (x wInterpolated), tan is not, vd
 * is normalized, the colour is weighted without 1/w (deviation D8). */
typedef struct {
    float uv[2], n[3], tan[3], vd[3], col[3];
} pixel_inputs;

static void interpolate_pixel(const scene_state *s, uint64_t g, uint32_t x, uint32_t y, pixel_inputs *pi,
                              int *mesh_out, uint32_t idx_out[3])
{
    uint32_t k;
    int m = locate_mesh(s, g, &k);
    const orc_mesh *mesh = &s->meshes[m];
    uint32_t idx[3];
    triangle_indices(mesh, k, idx);
    const tri_setup *t = &s->ts[g];
    float px = (float)x, py = (float)y;
    float w[3];
    w[2] = (px - t->sx[0]) * (t->sy[1] - t->sy[0]) - (py - t->sy[0]) * (t->sx[1] - t->sx[0]);
    w[0] = (px - t->sx[1]) * (t->sy[2] - t->sy[1]) - (py - t->sy[1]) * (t->sx[2] - t->sx[1]);
    w[1] = (px - t->sx[2]) * (t->sy[0] - t->sy[2]) - (py - t->sy[2]) * (t->sx[0] - t->sx[2]);
    w[0] /= t->area; w[1] /= t->area; w[2] /= t->area;
    const float *a0 = s->attr[m] + 16 * (size_t)idx[0];
    const float *a1 = s->attr[m] + 16 * (size_t)idx[1];
    const float *a2 = s->attr[m] + 16 * (size_t)idx[2];
    const float i0 = a0[14], i1 = a1[14], i2 = a2[14];
    const float wI = 1.f / (i0 * w[0] + i1 * w[1] + i2 * w[2]);
    float v[14];
    for (int c = 0; c < 14; ++c)
        v[c] = w[0] * (a0[c] * i0) + w[1] * (a1[c] * i1) + w[2] * (a2[c] * i2);
    pi->uv[0] = v[0] * wI; pi->uv[1] = v[1] * wI;
    pi->n[0] = v[2] * wI; pi->n[1] = v[3] * wI; pi->n[2] = v[4] * wI;
    pi->tan[0] = v[5]; pi->tan[1] = v[6]; pi->tan[2] = v[7];
    pi->vd[0] = v[8]; pi->vd[1] = v[9]; pi->vd[2] = v[10];
    normalize3(pi->vd);
    pi->col[0] = v[11]; pi->col[1] = v[12]; pi->col[2] = v[13];
    if (mesh_out) *mesh_out = m;
    if (idx_out) { idx_out[0] = idx[0]; idx_out[1] = idx[1]; idx_out[2] = idx[2]; }
}

static void *shade_worker(void *arg)
{
    shade_job *j = (shade_job *)arg;
    scene_state *s = j->s;
    const orc_opts *o = j->opts;
    const uint32_t W = s->W;
    for (uint32_t y = j->y0; y < j->y1; ++y) {
        for (uint32_t x = 0; x < W; ++x) {
            size_t pid = (size_t)x + (size_t)y * W;
            uint32_t g = j->vis[pid];
            if (g == ORC_VIS_NONE) {
                if (o->visualisation == ORC_VIS_PBR) j->rgba[pid] = o->clear_rgba;
                else j->rgba[pid] = 0xFF000000u;
                continue;
            }
            pixel_inputs pi;
            int m;
            uint32_t idx[3];
            interpolate_pixel(s, g, x, y, &pi, &m, idx);
            const orc_mesh *mesh = &s->meshes[m];
            if (j->vert_bitmap) {
                j->vert_bitmap[m][idx[0]] = 1; j->vert_bitmap[m][idx[1]] = 1; j->vert_bitmap[m][idx[2]] = 1;
            }
            float *uv = pi.uv, *n = pi.n, *tan = pi.tan, *vd = pi.vd, *col = pi.col;
            if (o->visualisation == ORC_VIS_DEPTH) { /* CCR/CUDARenderer.cu:862-879, Remap */
                float rgb[3] = { (j->depth[pid] - 0.985f) / (1.f - 0.985f), 0.f, 0.f };
                j->rgba[pid] = orc_pack_rgba(rgb);
                continue;
            }
            if (o->visualisation == ORC_VIS_NORMAL) { /* CCR/CUDARenderer.cu:882-897 */
                j->rgba[pid] = orc_pack_rgba(n);
                continue;
            }
            if (!mesh->tex[0].texels) { /* vertex colour, CCR/CUDARenderer.cu:855-858 */
                j->rgba[pid] = orc_pack_rgba(col);
                continue;
            }
            float smp[4][3];
            for (int slot = 0; slot < 4; ++slot) {
                const orc_texture *tx = mesh->tex[slot].texels ? &mesh->tex[slot] : &j->defaults[slot];
                uint8_t *bm = j->tex_bitmap ? j->tex_bitmap[m * 4 + slot] : NULL;
                sample_record(tx, uv, o->sample_state, smp[slot], bm);
            }
            float out[3];
            shade_pbr(smp[0], smp[1], smp[2], smp[3], n, tan, vd, out);
            j->rgba[pid] = orc_pack_rgba(out);
        }
    }
    return NULL;
}

int orc_render(const orc_mesh *meshes, int num_meshes, const orc_camera *cam, const orc_opts *opts,
               uint32_t W, uint32_t H, int nthreads, uint32_t *vis, float *depth, uint32_t *rgba,
               orc_stats *stats)
{
    if (!meshes || num_meshes <= 0 || !cam || !opts || W == 0 || H == 0 || !vis || !depth) return -1;
    if (nthreads < 1) nthreads = 1;
    scene_state s;
    if (scene_build(&s, meshes, num_meshes, cam, opts->cull_mode, W, H, nthreads) != 0) return -2;
    size_t P = (size_t)W * H;
    for (size_t i = 0; i < P; ++i) { vis[i] = ORC_VIS_NONE; depth[i] = 1.0f; }

    int nt = nthreads > (int)H ? (int)H : nthreads;
    raster_job *rj = (raster_job *)malloc(sizeof(raster_job) * (size_t)nt);
    if (!rj) { scene_free(&s); return -2; }
    for (int i = 0; i < nt; ++i) {
        rj[i].s = &s; rj[i].vis = vis; rj[i].depth = depth;
        rj[i].y0 = (uint32_t)((uint64_t)H * i / nt);
        rj[i].y1 = (uint32_t)((uint64_t)H * (i + 1) / nt);
    }
    run_parallel(raster_worker, rj, sizeof(raster_job), nt);
    free(rj);

    uint8_t **tex_bm = NULL, **vert_bm = NULL;
    orc_texture defaults[4];
    uint32_t dflt_texels[4] = { 0xFFFFFFFFu, opts->default_normal, opts->default_specular, opts->default_gloss };
    for (int i = 0; i < 4; ++i) { defaults[i].texels = &dflt_texels[i]; defaults[i].w = defaults[i].h = 1; }
    if (stats) {
        /* one bitmap per distinct texture object (texels pointer), defaults included */
        tex_bm = (uint8_t **)calloc((size_t)num_meshes * 4, sizeof(uint8_t *));
        vert_bm = (uint8_t **)calloc((size_t)num_meshes, sizeof(uint8_t *));
        if (!tex_bm || !vert_bm) { free(tex_bm); free(vert_bm); scene_free(&s); return -2; }
        for (int m = 0; m < num_meshes; ++m) {
            vert_bm[m] = (uint8_t *)calloc((size_t)meshes[m].num_verts + 1, 1);
            for (int slot = 0; slot < 4; ++slot) {
                const orc_texture *tx = meshes[m].tex[slot].texels ? &meshes[m].tex[slot] : &defaults[slot];
                uint8_t *found = NULL;
                for (int pm = 0; pm <= m && !found; ++pm)
                    for (int ps = 0; ps < 4; ++ps) {
                        if (pm == m && ps >= slot) break;
                        const orc_texture *ot = meshes[pm].tex[ps].texels ? &meshes[pm].tex[ps] : &defaults[ps];
                        if (ot->texels == tx->texels) { found = tex_bm[pm * 4 + ps]; break; }
                    }
                tex_bm[m * 4 + slot] = found ? found : (uint8_t *)calloc((size_t)tx->w * tx->h, 1);
            }
        }
    }
    if (rgba) {
        shade_job *sj = (shade_job *)malloc(sizeof(shade_job) * (size_t)nt);
        int n_shade = stats ? 1 : nt; /* bitmaps are written without locks */
        for (int i = 0; i < n_shade; ++i) {
            sj[i].s = &s; sj[i].opts = opts; sj[i].vis = vis; sj[i].depth = depth; sj[i].rgba = rgba;
            sj[i].tex_bitmap = tex_bm; sj[i].vert_bitmap = vert_bm; sj[i].defaults = defaults;
            sj[i].y0 = (uint32_t)((uint64_t)H * i / n_shade);
            sj[i].y1 = (uint32_t)((uint64_t)H * (i + 1) / n_shade);
        }
        run_parallel(shade_worker, sj, sizeof(shade_job), n_shade);
        free(sj);
    }
    if (stats) {
        memset(stats, 0, sizeof *stats);
        stats->num_triangles = s.num_tris;
        for (uint64_t g = 0; g < s.num_tris; ++g) stats->num_visible += s.alive[g];
        for (size_t i = 0; i < P; ++i) stats->num_covered_pixels += vis[i] != ORC_VIS_NONE;
        for (int m = 0; m < num_meshes; ++m) {
            for (uint32_t v = 0; v < meshes[m].num_verts; ++v) stats->distinct_ref_verts += vert_bm[m][v];
            for (int slot = 0; slot < 4; ++slot) {
                uint8_t *bm = tex_bm[m * 4 + slot];
                int first = 1;
                for (int q = 0; q < m * 4 + slot; ++q) if (tex_bm[q] == bm) { first = 0; break; }
                if (!first) continue;
                const orc_texture *tx = meshes[m].tex[slot].texels ? &meshes[m].tex[slot] : &defaults[slot];
                for (size_t q = 0; q < (size_t)tx->w * tx->h; ++q) stats->distinct_texels += bm[q];
            }
        }
        for (int q = 0; q < num_meshes * 4; ++q) {
            int first = 1;
            for (int p = 0; p < q; ++p) if (tex_bm[p] == tex_bm[q]) { first = 0; break; }
            if (first) free(tex_bm[q]);
        }
        for (int m = 0; m < num_meshes; ++m) free(vert_bm[m]);
        free(tex_bm); free(vert_bm);
    }
    scene_free(&s);
    return 0;
}

void orc_shade_pbr(const float diff[3], const float nrm[3], const float spc[3], const float gls[3],
                   const float n[3], const float tan[3], const float vd[3], float out[3])
{
    shade_pbr(diff, nrm, spc, gls, n, tan, vd, out);
}

int orc_shade_inputs(const orc_mesh *meshes, int num_meshes, const orc_camera *cam, int cull_mode,
                     uint32_t W, uint32_t H, uint32_t *vis, float *out)
{
    if (!meshes || num_meshes <= 0 || !cam || W == 0 || H == 0 || !vis || !out) return -1;
    scene_state s;
    if (scene_build(&s, meshes, num_meshes, cam, cull_mode, W, H, 1) != 0) return -2;
    const size_t P = (size_t)W * H;
    float *depth = (float *)malloc(P * sizeof(float));
    if (!depth) { scene_free(&s); return -2; }
    for (size_t i = 0; i < P; ++i) { vis[i] = ORC_VIS_NONE; depth[i] = 1.0f; }
    raster_job rj = { &s, 0, H, vis, depth };
    raster_worker(&rj);
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            const size_t pid = (size_t)x + (size_t)y * W;
            float *o = out + 11 * pid;
            if (vis[pid] == ORC_VIS_NONE) {
                for (int c = 0; c < 11; ++c) o[c] = NAN;
                continue;
            }
            pixel_inputs pi;
            interpolate_pixel(&s, vis[pid], x, y, &pi, NULL, NULL);
            o[0] = pi.uv[0]; o[1] = pi.uv[1];
            for (int c = 0; c < 3; ++c) { o[2 + c] = pi.n[c]; o[5 + c] = pi.tan[c]; o[8 + c] = pi.vd[c]; }
        }
    free(depth);
    scene_free(&s);
    return 0;
}

int orc_bin_counts(const orc_mesh *meshes, int num_meshes, const orc_camera *cam, int cull_mode,
                   uint32_t W, uint32_t H, uint32_t tw, uint32_t th, uint32_t big_threshold,
                   uint32_t *counts, uint32_t *big_ids, uint32_t big_cap, uint32_t *big_count)
{
    if (tw == 0 || th == 0) return -1;
    scene_state s;
    if (scene_build(&s, meshes, num_meshes, cam, cull_mode, W, H, 1) != 0) return -2;
    uint32_t ntx = (W + tw - 1) / tw, nty = (H + th - 1) / th;
    memset(counts, 0, sizeof(uint32_t) * (size_t)ntx * nty);
    uint32_t nb = 0;
    for (uint64_t g = 0; g < s.num_tris; ++g) {
        if (!s.alive[g]) continue;
        const tri_setup *t = &s.ts[g];
        if (t->xmin >= t->xmax || t->ymin >= t->ymax) continue;
        uint32_t tx0 = (uint32_t)t->xmin / tw, tx1 = (uint32_t)(t->xmax - 1) / tw;
        uint32_t ty0 = (uint32_t)t->ymin / th, ty1 = (uint32_t)(t->ymax - 1) / th;
        uint64_t span = (uint64_t)(tx1 - tx0 + 1) * (ty1 - ty0 + 1);
        if (span > big_threshold) {
            if (big_ids && nb < big_cap) big_ids[nb] = (uint32_t)g;
            ++nb;
            continue;
        }
        for (uint32_t ty = ty0; ty <= ty1; ++ty)
            for (uint32_t tx = tx0; tx <= tx1; ++tx) counts[ty * ntx + tx]++;
    }
    if (big_count) *big_count = nb;
    scene_free(&s);
    return 0;
}

int orc_render_bruteforce(const orc_mesh *meshes, int num_meshes, const orc_camera *cam, int cull_mode,
                          uint32_t W, uint32_t H, uint32_t *vis, float *depth)
{
    scene_state s;
    if (scene_build(&s, meshes, num_meshes, cam, cull_mode, W, H, 1) != 0) return -2;
    /* Per pixel, every triangle in ID order; the bbox only bounds the test
     * region (D4), the edge test and depth rule decide. */
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            float best = 1.0f;
            uint32_t id = ORC_VIS_NONE;
            for (uint64_t g = 0; g < s.num_tris; ++g) {
                if (!s.alive[g]) continue;
                const tri_setup *t = &s.ts[g];
                if ((int)x < t->xmin || (int)x >= t->xmax || (int)y < t->ymin || (int)y >= t->ymax) continue;
                /* GPU edge form (CCR/CUDAROPs.cu:120-145) on purpose: the
                 * sequential raster uses the CPU form, so agreement checks
                 * that the two forms are bitwise equal. */
                float px = (float)x, py = (float)y;
                float w2 = (px - t->sx[0]) * (t->sy[1] - t->sy[0]) - (py - t->sy[0]) * (t->sx[1] - t->sx[0]);
                float w0 = (px - t->sx[1]) * (t->sy[2] - t->sy[1]) - (py - t->sy[1]) * (t->sx[2] - t->sx[1]);
                float w1 = (px - t->sx[2]) * (t->sy[0] - t->sy[2]) - (py - t->sy[2]) * (t->sx[0] - t->sx[2]);
                if (!(w0 >= 0.f && w1 >= 0.f && w2 >= 0.f)) continue;
                float w[3] = { w0 / t->area, w1 / t->area, w2 / t->area };
                float z;
                if (!fragment_depth(t, w, &z)) continue;
                if (z < best || (z == best)) { best = z; id = (uint32_t)g; }
            }
            vis[(size_t)y * W + x] = id;
            depth[(size_t)y * W + x] = best;
        }
    scene_free(&s);
    return 0;
}

/* ------------------------------------------------------------------ MLP */

size_t orc_mlp_num_params(int in_dim, int hidden)
{
    return (size_t)hidden * in_dim + hidden + (size_t)hidden * hidden + hidden + hidden + 1;
}

/* Forward in double accumulation; the GPU uses fp32, compared with a tolerance. */
static void mlp_fwd1(const float *p, int in, int H, const float *x, double *h1, double *h2, double *out)
{
    const float *W1 = p, *b1 = W1 + (size_t)H * in, *W2 = b1 + H, *b2 = W2 + (size_t)H * H,
                *W3 = b2 + H, *b3 = W3 + H;
    for (int i = 0; i < H; ++i) {
        double a = b1[i];
        for (int k = 0; k < in; ++k) a += (double)W1[(size_t)i * in + k] * x[k];
        h1[i] = a > 0 ? a : 0;
    }
    for (int i = 0; i < H; ++i) {
        double a = b2[i];
        for (int k = 0; k < H; ++k) a += (double)W2[(size_t)i * H + k] * h1[k];
        h2[i] = a > 0 ? a : 0;
    }
    double o = b3[0];
    for (int k = 0; k < H; ++k) o += (double)W3[k] * h2[k];
    *out = o;
}

void orc_mlp_forward(const float *params, int in, int H, const float *x, int n, float *pred)
{
    double *h1 = (double *)malloc(sizeof(double) * (size_t)H * 2);
    if (!h1) return;
    for (int s = 0; s < n; ++s) {
        double o;
        mlp_fwd1(params, in, H, x + (size_t)s * in, h1, h1 + H, &o);
        pred[s] = (float)o;
    }
    free(h1);
}

double orc_mlp_loss_grad(const float *p, int in, int H, const float *x, const float *target, int n,
                         float tau, float *grad)
{
    size_t np = orc_mlp_num_params(in, H);
    double *g = (double *)calloc(np, sizeof(double));
    double *buf = (double *)malloc(sizeof(double) * (size_t)H * 4);
    if (!g || !buf) { free(g); free(buf); return NAN; }
    double *h1 = buf, *h2 = buf + H, *d2 = buf + 2 * H, *d1 = buf + 3 * H;
    const float *W2 = p + (size_t)H * in + H, *W3 = W2 + (size_t)H * H + H;
    double *gW1 = g, *gb1 = gW1 + (size_t)H * in, *gW2 = gb1 + H, *gb2 = gW2 + (size_t)H * H,
           *gW3 = gb2 + H, *gb3 = gW3 + H;
    double loss = 0;
    for (int s = 0; s < n; ++s) {
        const float *xs = x + (size_t)s * in;
        double o;
        mlp_fwd1(p, in, H, xs, h1, h2, &o);
        double r = (double)target[s] - o;
        loss += r > 0 ? tau * r : (tau - 1.0) * r;
        double dout = (r > 0 ? -(double)tau : 1.0 - tau) / n;
        gb3[0] += dout;
        for (int k = 0; k < H; ++k) {
            gW3[k] += dout * h2[k];
            d2[k] = h2[k] > 0 ? dout * W3[k] : 0;
        }
        for (int i = 0; i < H; ++i) {
            gb2[i] += d2[i];
            for (int k = 0; k < H; ++k) gW2[(size_t)i * H + k] += d2[i] * h1[k];
        }
        for (int k = 0; k < H; ++k) {
            double a = 0;
            for (int i = 0; i < H; ++i) a += (double)W2[(size_t)i * H + k] * d2[i];
            d1[k] = h1[k] > 0 ? a : 0;
        }
        for (int i = 0; i < H; ++i) {
            gb1[i] += d1[i];
            for (int k = 0; k < in; ++k) gW1[(size_t)i * in + k] += d1[i] * xs[k];
        }
    }
    for (size_t i = 0; i < np; ++i) grad[i] = (float)g[i];
    free(g); free(buf);
    return loss / n;
}
