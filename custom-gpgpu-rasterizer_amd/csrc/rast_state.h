// rast_state.h -- host-side objects behind the opaque handles of rast.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "rast.h"
#include "rast_internal.h"
#include "rast_launch.h"

// Records the calling thread's last error message (rast_last_error) and
// returns `code`.
int rast_set_error(int code, const char *fmt, ...);

#define HIP_TRY(expr)                                                                                      \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return rast_set_error(e_ == hipErrorOutOfMemory ? RAST_E_OUT_OF_MEMORY : RAST_E_HIP,           \
                                  "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__);     \
    } while (0)

#define RAST_TRY(expr)                  \
    do {                                \
        int rc_ = (expr);               \
        if (rc_ != RAST_OK) return rc_; \
    } while (0)

using rast::FrameBuffers;

// Drops cached batch graphs that refer to r (before r is destroyed).
void rast_forget_renderer(rast_renderer *r);

struct DBuf {
    void *p = nullptr;
    size_t n = 0;
};

inline void dfree(DBuf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
}

struct rast_texture {
    rast_renderer *r = nullptr;
    uint32_t *texels = nullptr;
    uint32_t w = 0, h = 0;
};

struct rast_mesh {
    rast_renderer *r = nullptr;
    float *vert16 = nullptr;  // 16 floats per vertex (IVertex + 2 pad floats)
    uint32_t *idx = nullptr;
    uint32_t nv = 0, ni = 0, ntri = 0;
    int topology = 0;
    float world[16];
    rast_texture *tex[4] = { nullptr, nullptr, nullptr, nullptr };
    uint64_t tri_base = 0;
};

struct FrameGraph {
    hipGraphExec_t exec = nullptr;
    uint64_t key = 0;
};

struct rast_renderer {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t W = 0, H = 0;
    uint32_t tile = 16, big_threshold = 256, small_limit = 16;
    std::vector<rast_mesh *> meshes;
    std::vector<rast_texture *> textures;
    rast_renderer *scene_src = nullptr;   // set for views: the renderer whose meshes they draw
    std::vector<rast_renderer *> views;    // views drawing this renderer's meshes
    uint64_t scene_version = 1;            // bumped when the mesh list changes
    uint64_t layout_version = 0;           // scene version the block tables were built for
    uint64_t generation = 0;               // bumped whenever a device buffer moves
    FrameGraph graphs[2];                  // captured frame per count-buffer parity
    hipEvent_t join = nullptr;             // batch capture: end of this renderer's branch
    rast::FrameParams fp{};
    FrameBuffers fb{};
    size_t staged_meshes = 0;
    // scene layout
    uint32_t total_verts = 0, total_tris = 0, vs_blocks = 0, tri_blocks = 0;
    std::vector<uint32_t> vs_block_mesh, tri_block_mesh;
    DBuf d_vs_block_mesh, d_tri_block_mesh, d_meshes;
    rast::DevMesh *h_meshes = nullptr;  // pinned staging
    size_t h_meshes_cap = 0;
    // frame buffers
    DBuf pos4, attr, trirec, alive, survivors;
    DBuf tile_count[2], tile_off, tile_cap, tile_pool, tile_cursor, density, tile_min_depth[2], features;
    DBuf queue, pool, big, ovf, counters, defaults;
    uint32_t queue_cap = 0, pool_cap = 0, big_cap = 0, ovf_cap = 0;
    DBuf vis, depth, rgba, keys;
    bool keys_dirty = true;  // key buffer must be reset before the next frame
    uint32_t micro_limit = 16;
    int cur = 0;  // index of the current frame's tile_count / tile_min_depth
    bool have_prev = false;
    rast::FrameCounters *h_counters = nullptr;  // pinned
    uint32_t *h_defaults = nullptr;       // pinned
    // timing
    hipEvent_t ev[12] = {};
    bool timing = false;
    int ev_used[12] = {};
    // last frame
    rast_camera last_cam{};
    rast_frame_opts last_opts{};
    bool frame_pending = false;
    bool frame_valid = false;
    bool last_single = false;
    bool prev_valid_before = false;
    bool frame_had_prev = false;  // the last frame had a previous frame (features valid)
    uint32_t rerenders = 0;
    rast_stats stats{};
    // learned capacities
    rast_binmodel *model = nullptr;
    float margin = 1.25f;
    uint32_t slack = 4;
};

