/*
 * rast_oracle.h -- CPU restatement of the reference rasterizer path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker / the timed CPU baseline.
 *
 * It restates the arithmetic of kdedeurw/Custom-GPGPU-Rasterizer (paths are
 * relative to that repository, CCR = CustomCUDARasterizer/):
 *   camera / matrices   CCR/Camera.cpp:26-31,97-117, CCR/Matrix4.h:120-166,327-363
 *   vertex transform    CCR/CUDAROPs.cu:99-118, CCR/ERenderer.cpp:215-242
 *   triangle assembly   CCR/ERenderer.cpp:73-157, CCR/CUDAROPs.cu:147-178
 *   bbox / raster       CCR/ERenderer.cpp:269-295,497-554,603-624,
 *                       CCR/CUDAROPs.cu:88-97,120-145,180-200,329-392
 *   pixel shading       CCR/CUDAROPs.cu:472-514, CCR/CUDARenderer.cu:833-897
 *   texture sampling    CCR/CUDATextureSampler.cuh:86-190, CCR/RGBRaw.h:23-38
 * with the documented deviations D1-D9 of the parity contract (see the
 * comments at each function).  Compiled with -ffp-contract=off and no
 * fast-math so that every float operation is one IEEE-754 binary32 op.
 */
#ifndef RAST_ORACLE_H
#define RAST_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Input vertex: CCR/Vertex.h:113-145 (IVertex, 56 bytes). */
typedef struct {
    float p[3];
    float uv[2];
    float n[3];
    float tan[3];
    float c[3];
} orc_ivertex;

/* RGBA8 texture, texel bytes in memory order R,G,B,A; texels == NULL = absent. */
typedef struct {
    const uint32_t *texels;
    uint32_t w, h;
} orc_texture;

enum { ORC_TOPO_LIST = 0, ORC_TOPO_STRIP = 1 };
enum { ORC_CULL_BACK = 0, ORC_CULL_FRONT = 1, ORC_CULL_NONE = 2 };
enum { ORC_SAMPLE_POINT = 0, ORC_SAMPLE_LINEAR = 1 };
enum { ORC_VIS_PBR = 0, ORC_VIS_DEPTH = 1, ORC_VIS_NORMAL = 2 };

typedef struct {
    const orc_ivertex *verts;
    uint32_t num_verts;
    const uint32_t *indices;
    uint32_t num_indices;
    int topology;
    float world[16];     /* reference layout: data[col][row] -> world[col*4+row] */
    orc_texture tex[4];  /* diffuse, normal, specular, gloss */
} orc_mesh;

typedef struct {
    float position[3];
    float forward[3];    /* reference Camera::m_Forward (opposite the view direction) */
    float fov_deg;
    float near_z, far_z;
} orc_camera;

typedef struct {
    int cull_mode;
    int sample_state;
    int visualisation;
    uint32_t clear_rgba;
    /* 1x1 stand-ins for missing normal/specular/gloss maps (deviation D8). */
    uint32_t default_normal, default_specular, default_gloss;
} orc_opts;

typedef struct {
    uint64_t num_triangles;       /* pre-cull, all meshes */
    uint64_t num_visible;         /* survivors of cull + frustum + degenerate test */
    uint64_t num_covered_pixels;  /* pixels with vis != NONE */
    uint64_t distinct_texels;     /* sum over texture objects of distinct texels fetched */
    uint64_t distinct_ref_verts;  /* distinct vertices of triangles that won >= 1 pixel */
} orc_stats;

/* Frame constants: CCR/CUDARenderer.cu:212-246 (host side of Render). */
typedef struct {
    float view[16], proj[16], viewproj[16];
    float cam_pos[3], cam_fwd[3], cam_right[3], cam_up[3];
} orc_frame;

#define ORC_VIS_NONE 0xFFFFFFFFu

void orc_camera_frame(const orc_camera *cam, uint32_t width, uint32_t height, orc_frame *out);
void orc_mat4_mul(const float *lm, const float *rm, float *out);
void orc_mat4_inverse(const float *m, float *out);

/* Per-vertex transform of one mesh: out_pos4[i] = (ndc.x, ndc.y, ndc.z, w),
 * out_attr[i] = {uv.x, uv.y, n'.xyz, tan'.xyz, vd.xyz, c.rgb, 1/w, 0}. */
void orc_vertex_transform(const orc_mesh *mesh, const orc_frame *fr,
                          float *out_pos4, float *out_attr16);

/* Triangle setup of all meshes in global ID order.  out_flags[t] = 1 if the
 * triangle survives.  out_rec[t*12..] = sx0,sy0,sx1,sy1,sx2,sy2,z0,z1,z2,area,
 * bbox(xmin|ymin<<16), bbox(xmax|ymax<<16) (as raw u32 bits). */
int orc_setup(const orc_mesh *meshes, int num_meshes, const orc_frame *fr,
              int cull_mode, uint32_t width, uint32_t height,
              uint8_t *out_flags, uint32_t *out_rec);

/* Whole frame: vis (u32 global triangle ID), depth (fp32), rgba (RGBA8). */
int orc_render(const orc_mesh *meshes, int num_meshes, const orc_camera *cam,
               const orc_opts *opts, uint32_t width, uint32_t height,
               int num_threads, uint32_t *vis, float *depth, uint32_t *rgba,
               orc_stats *stats);

/* Exact per-tile bin counts for tile_w x tile_h tiles.  Triangles whose tile
 * rectangle spans more than big_threshold tiles are not counted per tile but
 * listed in big_ids (ascending); big_count receives their number. */
int orc_bin_counts(const orc_mesh *meshes, int num_meshes, const orc_camera *cam,
                   int cull_mode, uint32_t width, uint32_t height,
                   uint32_t tile_w, uint32_t tile_h, uint32_t big_threshold,
                   uint32_t *counts, uint32_t *big_ids, uint32_t big_cap,
                   uint32_t *big_count);

/* Brute-force visibility: every pixel tested against every triangle with no
 * bounding box (pixels outside the bbox are still excluded, as in the spec).
 * Tiny inputs only. */
int orc_render_bruteforce(const orc_mesh *meshes, int num_meshes, const orc_camera *cam,
                          int cull_mode, uint32_t width, uint32_t height,
                          uint32_t *vis, float *depth);

/* ShadePixelUnsafe (CCR/CUDAROPs.cu:472-514) on given samples (rgb in [0,1])
 * and interpolated n, tan, vd: the shading step of orc_render. */
void orc_shade_pbr(const float diff[3], const float nrm[3], const float spc[3], const float gls[3],
                   const float n[3], const float tan[3], const float vd[3], float out[3]);

/* The shading inputs orc_render interpolates at every covered pixel: 11
 * floats per pixel (u, v, n.xyz, tan.xyz, vd.xyz), NaN where vis is none. */
int orc_shade_inputs(const orc_mesh *meshes, int num_meshes, const orc_camera *cam, int cull_mode,
                     uint32_t width, uint32_t height, uint32_t *vis, float *out);

/* Texture sampling (CCR/CUDATextureSampler.cuh:86-190): rgb in [0,1]. */
void orc_sample(const orc_texture *t, const float uv[2], int sample_state, float rgb[3]);

/* Colour packing with saturation (CCR/RGBRaw.h:23-32, CUDA cvt semantics). */
uint32_t orc_pack_rgba(const float rgb[3]);

/* ---- learned bin-capacity MLP (in -> H -> H -> 1, ReLU), fp32 ------------ */
/* Weight layout (row-major, out x in): W1[H*in], b1[H], W2[H*H], b2[H],
 * W3[H], b3[1].  Loss per sample: pinball(tau) on (target - pred), averaged. */
size_t orc_mlp_num_params(int in_dim, int hidden);
void orc_mlp_forward(const float *params, int in_dim, int hidden,
                     const float *x, int n, float *pred);
double orc_mlp_loss_grad(const float *params, int in_dim, int hidden,
                         const float *x, const float *target, int n, float tau,
                         float *grad);

#ifdef __cplusplus
}
#endif
#endif
