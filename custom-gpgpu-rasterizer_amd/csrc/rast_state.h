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

// Device tables, staging and (optionally) the captured graph of one launch
// shape: a renderer's single frames, or a cached batch shape.  Built for a
// structure key (frame sizes, tile sizes, binning modes, mesh sizes); the
// per-launch contents (cameras, options, buffer pointers) are restaged into
// the pinned block and uploaded by the launch itself.
struct LaunchPlan {
    uint64_t key = 0;
    bool built = false;
    rast::Launch L{};
    uint32_t njobs = 0, nmeshes = 0;
    std::vector<uint32_t> block0, nblocks;  // per job x stage (host copies for JobDesc)
    std::vector<uint32_t> vs_block0, tri_block0;  // per mesh entry
    DBuf d_block_mesh, d_tables, d_stage;
    DBuf d_alive_part;  // culling survivors per setup block
    // view-independent attributes of the scenes two or more jobs draw (views):
    // 32 bytes per vertex of each such scene, at sattr_off[job] (SIZE_MAX: none)
    DBuf d_sattr;
    std::vector<size_t> sattr_off;
    // Two pinned staging blocks (JobDesc[njobs] then DevMesh[nmeshes]) used in
    // turn, so the next launch is staged while the previous one still reads
    // its block: the host never waits for the frame before last.  One
    // captured graph per block (a graph's upload reads a fixed address).
    void *h_stage[2] = { nullptr, nullptr };
    size_t stage_bytes = 0;
    hipGraphExec_t exec[2] = { nullptr, nullptr };  // captured launches (graph mode)
    hipEvent_t done[2] = { nullptr, nullptr };      // end of the last launch that used each block
    bool done_pending[2] = { false, false };
    int parity = 0;                                 // the block the next launch uses
    // side streams running the resolve passes of different tile sizes concurrently
    hipStream_t side[2] = { nullptr, nullptr };
    hipEvent_t fork = nullptr, joined[2] = { nullptr, nullptr };
};
void plan_free(LaunchPlan &p);
// Waits until no queued GPU work uses r's buffers (its stream and, when its
// last frame came from a batch, that batch).
int rast_settle(rast_renderer *r);

struct rast_texture {
    rast_renderer *r = nullptr;
    uint32_t *texels = nullptr;
    uint32_t w = 0, h = 0;
};

// The four maps of a mesh interleaved (texel i = diffuse, normal, specular,
// gloss words of texel i), built when all four exist and share one size;
// shading then reads one 16-byte texel instead of four 4-byte ones.
struct TexQuad {
    rast_texture *t[4];
    uint4 *texels;
};

struct rast_mesh {
    rast_renderer *r = nullptr;
    float *ivert14 = nullptr;  // 14 floats per vertex (the 56-byte IVertex as given) + 16 bytes of padding
    uint32_t *idx = nullptr;
    uint32_t nv = 0, ni = 0, ntri = 0;
    int topology = 0;
    float world[16];
    rast_texture *tex[4] = { nullptr, nullptr, nullptr, nullptr };
    const uint4 *texq = nullptr;  // interleaved maps for tex[0..3], set by prepare_job
    float attr_max = 0.f;         // largest |component| of the input normals and tangents
    uint64_t tri_base = 0;
};

struct rast_renderer {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t W = 0, H = 0;
    uint32_t tile = 16, big_threshold = 256, small_limit = 16;
    int resolve_mode = 0;                  // RAST_RESOLVE_*
    std::vector<rast_mesh *> meshes;
    std::vector<rast_texture *> textures;
    std::vector<TexQuad> tex_quads;        // interleaved map sets of this renderer's textures
    rast_renderer *scene_src = nullptr;   // set for views: the renderer whose meshes they draw
    std::vector<rast_renderer *> views;    // views drawing this renderer's meshes
    uint64_t scene_version = 1;            // bumped when the mesh list changes
    uint64_t layout_version = 0;           // scene version the block tables were built for
    hipEvent_t join = nullptr;             // orders a batch after this renderer's earlier work
    uint64_t batch_seq = 0;                // batch that rendered the last frame (0: none pending)
    bool stream_busy = false;              // work queued on `stream` since its last synchronization
    LaunchPlan plan;                       // single-frame launches
    // scene layout
    uint32_t total_verts = 0, total_tris = 0;
    // frame buffers
    DBuf pos4, attr, trirec, alive, survivors;
    DBuf tile_count[2], tile_off, tile_cap, tile_pool, tile_cursor, tile_min_depth[2];
    DBuf queue, pool, big, bigp, ovf;
    uint32_t queue_cap = 0, pool_cap = 0, big_cap = 0, ovf_cap = 0;
    DBuf vis, depth, rgba, keys;
    bool keys_dirty = true;  // key buffer must be reset before the next frame
    uint32_t micro_limit = 16;
    int cur = 0;  // index of the current frame's tile_count / tile_min_depth
    bool have_prev = false;
    rast::FrameCounters *h_counters = nullptr;  // pinned, written by the launch
    rast::FrameCounters *d_counters_host = nullptr;  // device address of h_counters
    // timing
    hipEvent_t ev[12] = {};
    bool timing = false;
    bool timer_started = false;  // rast_timer_start recorded its event
    int ev_used[12] = {};
    // last frame
    rast_camera last_cam{};
    rast_frame_opts last_opts{};
    bool frame_pending = false;
    bool frame_valid = false;
    bool last_single = false;
    bool last_shared_attr = false;  // the last frame's attributes were split (views rendered together)
    bool prev_valid_before = false;
    bool frame_had_prev = false;  // the last frame had a previous frame (features valid)
    uint32_t rerenders = 0;
    rast_stats stats{};
    // learned capacities
    rast_binmodel *model = nullptr;
    float margin = 1.25f;
    uint32_t slack = 4;
};

