/*
 * rast.h -- C ABI of the MI355X (gfx950) tiled rasterizer.
 *
 * This is the drop-in boundary for the GPU rendering path of
 * kdedeurw/Custom-GPGPU-Rasterizer (paths below are relative to that
 * repository; CCR = CustomCUDARasterizer/).  Each entry point names the
 * reference interface it replaces.  Plain pointers, sizes and POD structs
 * only; every call returns RAST_OK (0) or a negative RAST_E_* code, and
 * rast_last_error() returns the message of the calling thread's last failure.
 * No call throws.  A renderer owns one HIP stream; rendering is asynchronous
 * until rast_sync / rast_readback.  Handles are not thread-safe.
 *
 * Output contract of a frame (deterministic, bit-reproducible):
 *   vis   u32 per pixel: global triangle ID (mesh base + triangle index in
 *         index-buffer order, before culling) of the visible triangle, or
 *         RAST_VIS_NONE;
 *   depth f32 per pixel: NDC z of that triangle at the pixel, 1.0 if none;
 *   rgba  RGBA8 per pixel (memory order R,G,B,A).
 * A fragment at integer pixel (x, y) exists where all three edge functions are
 * >= 0 and 0 <= z <= 1; the smallest z wins and, on equal z, the larger ID
 * (the sequential "later triangle wins ties" rule of CCR/ERenderer.cpp:544-554).
 */
#ifndef RAST_H
#define RAST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RAST_OK 0
#define RAST_E_INVALID_ARG (-1)
#define RAST_E_HIP (-2)
#define RAST_E_OUT_OF_MEMORY (-3)
#define RAST_E_STATE (-4)
#define RAST_E_NO_DEVICE (-5)

#define RAST_VIS_NONE 0xFFFFFFFFu

typedef struct rast_renderer rast_renderer;
typedef struct rast_mesh rast_mesh;
typedef struct rast_texture rast_texture;
typedef struct rast_binmodel rast_binmodel;

/* Input vertex, 56 bytes: CCR/Vertex.h:113-145 (IVertex). */
typedef struct {
    float p[3];
    float uv[2];
    float n[3];
    float tan[3];
    float c[3];
} rast_ivertex;

/* CCR/PrimitiveTopology.h:3-7 */
enum { RAST_TOPOLOGY_LIST = 0, RAST_TOPOLOGY_STRIP = 1 };
/* CCR/CullingMode.h:3-8 */
enum { RAST_CULL_BACK = 0, RAST_CULL_FRONT = 1, RAST_CULL_NONE = 2 };
/* CCR/SampleState.h:3-7 */
enum { RAST_SAMPLE_POINT = 0, RAST_SAMPLE_LINEAR = 1 };
/* CCR/VisualisationState.h:3-8 */
enum { RAST_VIS_PBR = 0, RAST_VIS_DEPTH = 1, RAST_VIS_NORMAL = 2 };
/* Bin-queue policy: exact two-pass (count, scan, fill); single pass with
 * capacities from the learned model; single pass with previous-frame
 * counts x margin.  The single-pass policies fall back to exact binning for
 * every tile whose capacity overflowed, so the frame output never changes. */
enum { RAST_BIN_EXACT = 0, RAST_BIN_LEARNED = 1, RAST_BIN_PREVFRAME = 2 };

/* Camera of CCR/Camera.h:6-31 (position, m_Forward, FOV degrees, near, far).
 * `forward` is the reference's m_Forward: it points away from the view
 * direction (the default camera at (0,0,3) with forward (0,0,1) looks at the
 * origin). */
typedef struct {
    float position[3];
    float forward[3];
    float fov_deg;
    float near_z;
    float far_z;
} rast_camera;

/* Per-frame options: CCR/SceneManager.cpp:59-144 runtime toggles plus the
 * binner policy.  Missing normal / specular / gloss maps are replaced by 1x1
 * textures of the given RGBA8 values; a mesh without a diffuse map is shaded
 * with its interpolated vertex colour (CCR/CUDARenderer.cu:848-858). */
typedef struct {
    int cull_mode;
    int sample_state;
    int visualisation;
    int bin_policy;
    uint32_t clear_rgba;
    uint32_t default_normal;
    uint32_t default_specular;
    uint32_t default_gloss;
    uint32_t flags;          /* RAST_FLAG_* */
} rast_frame_opts;

#define RAST_FLAG_TIMING 1u   /* record per-stage HIP events (read with rast_get_stats) */
#define RAST_FLAG_NO_SHADE 2u /* stop after visibility + depth */
/* Replay the frame from a captured HIP graph (captured on first use, kept per
 * options, sizes and tuning, re-captured when a device buffer moves).  Has no
 * reference counterpart: it replaces the per-stage launches of
 * CCR/CUDARenderer.cu:237-330 with one graph launch.  Ignored with TIMING. */
#define RAST_FLAG_GRAPH 4u

/* Frame statistics: CCR/CUDARenderer.h:63-64 (visible / total triangles) and
 * the per-stage milliseconds of the BENCHMARK build (CCR/CUDARenderer.cu:344-353). */
typedef struct {
    uint64_t num_triangles;
    uint64_t num_visible;        /* survivors of cull, frustum and degenerate tests */
    uint64_t num_bin_pairs;      /* (tile, triangle) entries written to the per-tile queues */
    uint32_t num_big;            /* triangles kept on the screen-wide list */
    uint32_t num_tiles;
    uint32_t overflow_tiles;     /* single-pass policies: tiles that fell back to exact binning */
    uint32_t capacity_total;     /* single-pass policies: entries allocated by the capacity model */
    uint32_t rerenders;          /* frames re-run after a queue-storage overflow (storage grown) */
    uint32_t overflow_entries;   /* single-pass policies: queue entries placed by the exact fallback */
    float ms_vertex, ms_setup, ms_count, ms_scan, ms_fill, ms_fallback, ms_raster, ms_shade, ms_total;
    float ms_model;              /* single pass: tile capacities (model or heuristic) and queue regions */
} rast_stats;

/* Device query: CCR/CUDARenderer.cu:59-175 (DisplayGPUSpecs). */
typedef struct {
    char name[256];
    char arch[64];
    int compute_units;
    int clock_khz;
    uint64_t global_mem_bytes;
    int lds_bytes_per_block;
    int wavefront_size;
} rast_device_info;

/* Frame constants produced by the host camera math: CCR/Camera.cpp:97-117,
 * CCR/Camera.h:140-143, CCR/CUDARenderer.cu:212-217.  Matrices use the
 * reference's storage data[col][row] (m[col * 4 + row]). */
typedef struct {
    float view[16], proj[16], viewproj[16];
    float cam_pos[3], cam_fwd[3], cam_right[3], cam_up[3];
} rast_frame_consts;

/* ---------------------------------------------------------------- errors */
/* Replaces CheckErrorCuda / GPUAssert printing (CCR/PCH.h:46-56,
 * CCR/GPUHelpers.h:16-25): copies the last error message, returns its length. */
int rast_last_error(char *buf, size_t buf_len);
int rast_device_query(int device, rast_device_info *out);

/* -------------------------------------------------------------- renderer */
/* CUDARenderer(const WindowHelper&) + AllocateCUDADeviceBuffers
 * (CCR/CUDARenderer.h:34, CCR/CUDARenderer.cu:32-45,413-475): a renderer for a
 * width x height framebuffer on HIP device `device`. */
int rast_renderer_create(int device, uint32_t width, uint32_t height, rast_renderer **out);
/* A second framebuffer over the meshes of `owner` (no reference counterpart:
 * the reference renders one view per renderer; cf. LoadScene
 * CCR/CUDARenderer.cu:57-75).  Views own their frame buffers, bins, capacity
 * model and captured graphs; meshes and textures stay with the owner, which
 * cannot be destroyed while views exist.  Creating meshes on a view fails. */
int rast_renderer_create_view(rast_renderer *owner, uint32_t width, uint32_t height, rast_renderer **out);
/* ~CUDARenderer (CCR/CUDARenderer.cu:47-52): waits for the stream, frees the
 * renderer's buffers and the device memory of all its meshes and textures.
 * Their handles stay allocated, empty: rast_mesh_destroy / rast_texture_destroy
 * on them only frees the handle (so the objects holding them may be destroyed
 * after the renderer, the order of CCR/main.cpp:418,529), and the other calls
 * taking them return RAST_E_STATE or RAST_E_INVALID_ARG. */
int rast_renderer_destroy(rast_renderer *r);
/* SetupBins(numBins, binDim, binQueueMaxSize) (CCR/CUDARenderer.cu:199-204):
 * square tiles of tile_size pixels (8, 16 or 32); triangles whose tile
 * rectangle spans more than big_threshold tiles go to one screen-wide list
 * instead of the per-tile queues (0 = default 256). */
int rast_setup_bins(rast_renderer *r, uint32_t tile_size, uint32_t big_threshold);
/* Raster tuning (no reference counterpart; the reference derives its
 * per-thread pixel coverage from the bin size, CCR/main.cpp:372-389):
 * triangles whose bounding box has at most micro_limit pixels are rasterized
 * directly at setup with 64-bit atomic min (0 = bin every triangle); inside a
 * tile, triangles covering at most small_limit pixels of it are rasterized one
 * per lane, larger ones one pixel per lane.  Defaults 16 / 16.  The output
 * does not depend on either value. */
int rast_set_tuning(rast_renderer *r, uint32_t micro_limit, uint32_t small_limit);
/* Resolve-stage launch form (the persistent kernel of the reference's TODO,
 * CCR/CUDARenderer.cu:284-285): RAST_RESOLVE_PER_TILE (default) one
 * workgroup per tile; RAST_RESOLVE_PERSISTENT workgroups as many as the GPU
 * holds, each taking the next tile of any job from one counter;
 * RAST_RESOLVE_PERSISTENT_SHADE the same, each tile shaded by the workgroup
 * that resolved it (when the launch's shaded frames share one sampling mode;
 * otherwise shading stays a pass of its own).  A batch takes its first
 * job's form.  The output does not depend on it. */
enum { RAST_RESOLVE_PER_TILE = 0, RAST_RESOLVE_PERSISTENT = 1, RAST_RESOLVE_PERSISTENT_SHADE = 2 };
int rast_set_resolve_mode(rast_renderer *r, int mode);
/* Resize the framebuffer (a new Resolution in CCR/Resolution.cpp). */
int rast_resize(rast_renderer *r, uint32_t width, uint32_t height);

/* ------------------------------------------------------------------ mesh */
/* CUDAMesh(const Mesh*) + LoadScene (CCR/CUDAMesh.cpp:4-20,42-59,
 * CCR/CUDARenderer.cu:177-197): copies the vertices and indices to device
 * memory.  Meshes are drawn in creation order; a mesh's triangles take the
 * next block of global triangle IDs.  `world` uses data[col][row] storage
 * (NULL = identity). */
int rast_mesh_create(rast_renderer *r, const rast_ivertex *verts, uint32_t num_verts,
                     const uint32_t *indices, uint32_t num_indices, int topology,
                     const float world[16], rast_mesh **out);
int rast_mesh_destroy(rast_mesh *m);
/* CUDAMesh::GetWorld() (CCR/CUDAMesh.h:31). */
int rast_mesh_set_world(rast_mesh *m, const float world[16]);
/* CUDAMesh::SetTextures / CUDARenderer::UpdateCUDAMeshTextures
 * (CCR/CUDAMesh.h:42, CCR/CUDARenderer.cu:534-540): diffuse, normal,
 * specular, gloss; NULL = absent. */
int rast_mesh_set_textures(rast_mesh *m, rast_texture *diffuse, rast_texture *normal,
                           rast_texture *specular, rast_texture *gloss);
/* CUDAMesh::GetTotalNumTriangles (CCR/CUDAMesh.cpp:27-40). */
int rast_mesh_num_triangles(const rast_mesh *m, uint32_t *out);
/* Base global triangle ID of the mesh in the current scene order. */
int rast_mesh_triangle_base(const rast_mesh *m, uint64_t *out);

/* --------------------------------------------------------------- texture */
/* CUDATexture::Create (CCR/CUDATexture.cpp:24-72) from RGBA8 texels already in
 * memory (row-major, pitch = width) instead of an image path. */
int rast_texture_create(rast_renderer *r, const uint32_t *rgba8, uint32_t width, uint32_t height,
                        rast_texture **out);
/* The same from rows pitch_bytes apart (an image surface's pitch, as
 * CCR/CUDATexture.cpp:36-45 copies with cudaMemcpy2D); RGBA8 only.  The
 * texels are stored packed. */
int rast_texture_create_pitched(rast_renderer *r, const void *rgba8, uint32_t width, uint32_t height,
                                size_t pitch_bytes, rast_texture **out);
/* CUDATexture::Destroy (CCR/CUDATexture.cpp:74-82).  Meshes using it fall
 * back to "absent" for that slot. */
int rast_texture_destroy(rast_texture *t);

/* ------------------------------------------------------------------ frame */
/* Host camera math of CCR/Camera.cpp:26-31,97-117 (exposed for tests). */
int rast_camera_frame(const rast_camera *cam, uint32_t width, uint32_t height, rast_frame_consts *out);
/* CUDARenderer::Render (CCR/CUDARenderer.cu:206-354) including the buffer
 * clear of EnterValidRenderingState (CCR/CUDARenderer.cu:373-381): vertex
 * transform, triangle setup/cull, tile binning, per-tile raster with depth
 * test and deferred shading, enqueued on the renderer's stream. */
int rast_render(rast_renderer *r, const rast_camera *cam, const rast_frame_opts *opts);
/* Waits for the renderer's stream (CUDARenderer's cudaDeviceSynchronize). */
int rast_sync(rast_renderer *r);
/* CUDARenderer::Present (CCR/CUDARenderer.cu:383-395): copies the frame to
 * host memory; any pointer may be NULL.  Waits for the frame. */
int rast_readback(rast_renderer *r, uint32_t *vis, float *depth, uint32_t *rgba);
/* GetTotalNumVisibleTriangles / GetTotalNumTriangles + BENCHMARK timings. */
int rast_get_stats(rast_renderer *r, rast_stats *out);
/* Device pointers of the current frame's vis / depth / rgba buffers, for
 * zero-copy consumers.  Waits for the frame and checks it (a frame whose
 * queue storage overflowed is rendered again first), as rast_readback does;
 * RAST_E_STATE if no frame was rendered.  The pointers stay valid until the
 * next rast_render / rast_render_batch / rast_resize of the renderer. */
int rast_frame_buffers(rast_renderer *r, void **vis, void **depth, void **rgba);
/* Test-only inspection of intermediate buffers (the reference's test kernels
 * DrawTexture/DrawTextureGlobal play this role, CCR/CUDARenderer.h:42-45):
 * which = RAST_DEBUG_*; copies at most `bytes` and stores the size. */
enum {
    RAST_DEBUG_VERTEX_POS = 0,  /* float4 per vertex (ndc x, y, z, clip w), all meshes */
    RAST_DEBUG_VERTEX_ATTR = 1, /* 12 floats per vertex, each attribute times 1/w: u v 1/w n.x | n.y n.z t.x t.y
                                   | t.z vd.xyz (the vertex colour is applied at shading); RAST_E_STATE after a
                                   batch in which other views drew this renderer's scene (their
                                   attributes are stored split, the view-independent part once) */
    RAST_DEBUG_TRI_RECORDS = 2, /* 16 words per triangle (valid for survivors): sx0 sy0 sx1 sy1 sx2 sy2 z0 area
                                   z2 z1 bbox-lo bbox-hi i0 i1 i2 mesh */
    RAST_DEBUG_TRI_ALIVE = 3,   /* u8 per triangle */
    RAST_DEBUG_TILE_COUNTS = 4, /* u32 per tile: queue length */
    RAST_DEBUG_BIG_LIST = 5,    /* u32 count followed by the IDs */
    RAST_DEBUG_TILE_CAPS = 6    /* u32 per tile: single-pass queue capacity of the last frame */
};
int rast_debug_read(rast_renderer *r, int which, void *dst, size_t bytes, size_t *size_out);

/* CUDARenderer::StartTimer / StopTimer and GetBenchMarker's CUDABenchMarker
 * (CCR/CUDARenderer.cu:398-412, CCR/CUDABenchMarker.cpp:24-35): HIP events on
 * the renderer's stream.  rast_timer_stop records the stop event (after the
 * renderer's last batch, if its frame came from one), waits for it and stores
 * the milliseconds since rast_timer_start's event. */
int rast_timer_start(rast_renderer *r);
int rast_timer_stop(rast_renderer *r, float *ms);

/* Test hooks DrawTexture / DrawTextureGlobal (CCR/CUDARenderer.cu:1969-2012,
 * kernels :1565-1617): the renderer's frame becomes one texture (vis
 * RAST_VIS_NONE, depth 1.0 everywhere), readable with rast_readback.
 *   RAST_DRAW_TEXTURE: pixel (x, y) takes the texel nearest to the normalized
 *     coordinates (x / width, y / height) under wrap addressing, texel
 *     (floor(u * w) mod w, floor(v * h) mod h), copied unchanged (the
 *     reference's tex2D point fetch; its R/B swap converts the texture's byte
 *     order to its window surface's, and both are R,G,B,A here);
 *   RAST_DRAW_TEXTURE_GLOBAL: the colour CUDATextureSampler::Sample returns
 *     for uv = (x, y) / (width, height) (stretch != 0) or (x, y) / (w, h),
 *     in sample_state's mode (magenta outside [0, 1]), packed as shading packs
 *     colours.  sample_state is ignored by RAST_DRAW_TEXTURE. */
enum { RAST_DRAW_TEXTURE = 0, RAST_DRAW_TEXTURE_GLOBAL = 1 };
int rast_draw_texture(rast_renderer *r, const rast_texture *t, int mode, int stretch, int sample_state);

/* --------------------------------------------------------- learned binner */
/* No reference counterpart (the reference sizes queues statically,
 * CCR/DEFINES.h:21-23, CCR/main.cpp:372-389): a per-tile MLP in -> H -> H -> 1
 * that predicts log(1 + queue length) from per-tile features (previous-frame
 * queue lengths and min depth, current projected-vertex density), trained on
 * the GPU from exact counts. */
int rast_binmodel_create(rast_renderer *r, uint32_t hidden, uint32_t seed, rast_binmodel **out);
int rast_binmodel_destroy(rast_binmodel *m);
int rast_binmodel_num_features(void);
/* Attach (or detach with NULL) the model used by RAST_BIN_LEARNED.  A tile's
 * queue capacity is ceil(expm1(prediction) * margin) + slack (previous-frame
 * policy: ceil(count * margin) + slack), at most the frame's triangle count;
 * margin in [1, 64], slack at most 2^20.  Capacities are hints: the queue
 * storage bounds their sum, and entries beyond a tile's capacity are placed
 * by the exact fallback. */
int rast_renderer_set_binmodel(rast_renderer *r, rast_binmodel *m, float margin, uint32_t slack);
/* Per-tile features and exact log-count targets of the last rendered frame,
 * appended to the model's device-side training set. */
int rast_binmodel_record(rast_binmodel *m, rast_renderer *r);
int rast_binmodel_clear_data(rast_binmodel *m);
/* Adam on the pinball loss (quantile tau) over the recorded set. */
int rast_binmodel_train(rast_binmodel *m, uint32_t epochs, uint32_t batch, float lr, float tau,
                        float *final_loss);
int rast_binmodel_get_params(rast_binmodel *m, float *dst, size_t count);
int rast_binmodel_set_params(rast_binmodel *m, const float *src, size_t count);
int rast_binmodel_num_params(rast_binmodel *m, size_t *out);
/* Test hooks: one forward / loss-gradient evaluation on host-provided data. */
int rast_binmodel_forward(rast_binmodel *m, const float *x, uint32_t n, float *pred);
/* No counterpart in the reference.  rast_binmodel_forward for callers that
 * cannot allocate (serving loops, code that must not call hipMalloc): the
 * same kernel and result, with its device staging in a static array of the
 * library (2.25 MiB on every GPU it runs on, held until the process ends)
 * instead of buffers the model grows.  At most 65536 samples per call
 * (RAST_E_INVALID_ARG above that).  Calls on one GPU share that array; the
 * library serializes them (a call returns after its predictions are copied
 * out), so overlapping calls from several threads wait for each other. */
int rast_binmodel_forward_noalloc(rast_binmodel *m, const float *x, uint32_t n, float *pred);
int rast_binmodel_loss_grad(rast_binmodel *m, const float *x, const float *target, uint32_t n, float tau,
                            float *loss, float *grad);

/* ----------------------------------------------------- batched rendering */
/* One render job of a batch: no reference counterpart; it realises the
 * "persistent kernel approach" TODO of CCR/CUDARenderer.cu:284-286 and the
 * per-mesh host loop of CCR/CUDARenderer.cu:237-330.  rast_render_batch
 * enqueues all jobs as one captured HIP graph in which every pipeline stage
 * is one kernel over the blocks of all jobs (stage-fused, so small frames
 * share the chip); the graph is cached per batch structure (sizes, tile
 * sizes, binning modes) and replayed with the new cameras and options.
 * Every job needs its own renderer (views of one scene are fine); timing
 * flags are ignored.  Results are read per renderer with rast_readback /
 * rast_get_stats as after rast_render. */
typedef struct {
    rast_renderer *target;   /* renderer providing the framebuffer and meshes */
    rast_camera camera;
    rast_frame_opts opts;
} rast_job;
int rast_render_batch(const rast_job *jobs, uint32_t num_jobs);

#ifdef __cplusplus
}
#endif
#endif
