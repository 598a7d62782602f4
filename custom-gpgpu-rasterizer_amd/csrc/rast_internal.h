// rast_internal.h -- device data layout shared by the kernels and the host
// side of librast.so.  gfx950 (CDNA4, wave64) only.
//
// HBM layout of one frame (all device buffers owned by the renderer):
//   ivert    per mesh, 14 floats / vertex (the 56-byte IVertex, as uploaded)
//   pos4     float4 / global vertex: NDC x, y, z and clip w        (16 B)
//   attr     4 x float4 / global vertex, each attribute times 1/w:
//            u v 1/w n'.x | n'.yz tan'.xy | tan'.z vd | c.rgb 0
//            (64 B; PBR shading reads the first 48 B)
//   trirec   16 words / global triangle ID (written for survivors):
//            [0..5] screen x0,y0,x1,y1,x2,y2  [6..8] z0,z1,z2  [9] |area|
//            [10] xmin | ymin << 16  [11] xmax | ymax << 16 (pixel range
//            [min, max)), [12..14] global vertex indices, [15] mesh index,
//            bit 31 set where shading must follow the reference's operation
//            order (a vertex near or behind the camera plane, or a mesh with
//            large normals / tangents: the fused forms could drift there)
//   alive    u8 / triangle
//   tiles    count / offset / cursor / capacity u32 per tile, queue entries
//            u32 triangle IDs, plus one screen-wide "big" list
//   keys     u64 per pixel: (z bits << 32) | ~(id + 1), min = visible
//            fragment; reset to (1.0, none) by the resolve pass
//   vis, depth  u32 / f32 per pixel, rgba u32 per pixel
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rast {

constexpr uint32_t kVisNone = 0xFFFFFFFFu;
constexpr int kIVertexFloats = 14;
// Setup record, 16 u32 per triangle: sx0 sy0 sx1 sy1 | sx2 sy2 z0 area |
// z2 z1 bbox-lo bbox-hi | i0 i1 i2 mesh (area next to the screen positions,
// so shading reads quads 0, 1 and 3 only)
constexpr int kTriRecWords = 16;
constexpr int kMaxMeshes = 1 << 16;
constexpr int kNumFeatures = 8;
// Tiles per capacity workgroup (one per lane of its four waves).
constexpr int kCapacityTilesPerBlock = 256;
constexpr int kShadeRows = 8;  // pixel rows per shading block (64 wide; a multiple of 4)
constexpr uint32_t kVertexChunk = 128;  // vertices per vertex workgroup (256 threads; at most 256)

struct DevTexture {
    const uint32_t *texels;
    uint32_t w, h;
};

// One mesh of the frame, rebuilt on the host each frame (matrices change).
struct DevMesh {
    const float *ivert;       // 14 floats per vertex
    const uint32_t *idx;
    uint32_t num_verts;
    uint32_t num_tris;
    uint32_t vert_base;       // first global vertex index
    uint32_t tri_base;        // first global triangle ID
    uint32_t vs_block0;       // first block of this mesh's vertex group in the vertex launch
    uint32_t tri_block0;      // first block of this mesh in the triangle launch
    int32_t topology;
    int32_t has_diffuse;
    uint32_t job;             // index of the frame (job) this mesh entry belongs to
    uint32_t shade_exact;     // 1: the mesh's normals / tangents (times the world matrix) may exceed 4
                              // in magnitude; its triangles are shaded in the reference's operation order
    float wvp[16];            // data[col][row]
    float world[16];
    float cam_pos[4];         // camera position (xyz) and forward (xyz) of the frame
    float cam_fwd[4];
    DevTexture tex[4];        // diffuse, normal, specular, gloss (stand-ins filled in)
    const uint4 *texq;        // the four maps interleaved, one 16-byte texel each (null unless one size)
};

// Per-frame values that do not depend on the camera (the camera travels in
// DevMesh), so a captured frame graph stays valid while the camera moves.
struct FrameParams {
    uint32_t width, height;
    int32_t cull_mode;
    int32_t sample_state;
    int32_t visualisation;
    uint32_t clear_rgba;
    uint32_t tile;            // tile edge in pixels
    uint32_t tiles_x, tiles_y;
    uint32_t big_threshold;   // max tile span kept in per-tile queues
    uint32_t micro_limit;     // bbox pixel count up to which setup rasterizes directly
    uint32_t num_tris;        // all meshes
    uint32_t num_verts;
    uint32_t num_meshes;
};

// Launch stages whose grids are shared by all jobs of a launch: block b of
// stage s serves job table_s[b], as its block b - JobDesc::block0[s].
enum Stage : int {
    ST_INIT,      // per-tile reset (256 tiles per block)
    ST_CAP,       // single-pass capacities and queue regions (256 tiles per block)
    ST_FILL,      // exact second pass (grid-stride over binned triangles)
    ST_FBALLOC,   // previous-frame fallback allocation (256 tiles per block)
    ST_FBFILL,    // previous-frame fallback placement (grid-stride over entries)
    ST_RES8,      // resolve, 8x8 tiles (one block per tile, in launch order)
    ST_RES16,     // resolve, 16x16 tiles
    ST_RES32,     // resolve, 32x32 tiles
    ST_SHADE,     // 64 x kShadeRows-pixel blocks
    ST_COUNT
};

// Device-side frame counters, read back once per frame.
struct FrameCounters {
    uint32_t num_alive;       // survivors of culling
    uint32_t num_binned;      // survivors with a non-empty pixel range and span <= big_threshold
    uint32_t num_big;
    uint32_t total_pairs;     // exact queue entries (sum of per-tile counts)
    uint32_t overflow_tiles;  // single pass: tiles whose count exceeded capacity
    uint32_t ovf_entries;     // single pass: queue entries beyond capacity (fallback input)
    uint32_t pool_used;       // fallback pool entries
    uint32_t storage_overflow;// queue / pool / big list storage exhausted -> frame invalid
    uint32_t capacity_total;  // single pass: capacities as used (their sum fits the queue storage)
    uint32_t capacity_wanted; // single pass: the capacities' sum before that bound (saturated)
    unsigned long long capacity_top;  // single pass: queue allocation top (the capacities' full sum);
                                      // the two fields above are derived from it when published
};

// One frame of a launch: its parameters and the device buffers of the
// renderer that owns it.  Uploaded with the mesh table for every launch, so
// a captured launch graph replays with new cameras, options and buffer
// parities; only the grid shapes are fixed by the graph.
struct alignas(16) JobDesc {  // 16-byte multiple: staged in 16-byte words
    FrameParams fp;
    uint32_t block0[ST_COUNT];   // first block of this job in each stage's grid
    uint32_t nblocks[ST_COUNT];  // its block count there (0: not part of that stage)
    uint32_t mesh0;              // first entry of the job in the mesh table
    uint32_t single;             // single-pass binning (capacities known before setup)
    uint32_t learned;            // capacities from the MLP (else previous count * margin + slack)
    uint32_t small_limit;        // tile-clipped pixel count from which resolve goes pixel-parallel
    float margin;
    uint32_t slack;
    uint32_t queue_cap, pool_cap, big_cap, ovf_cap;
    uint32_t defaults[4];        // 1x1 stand-in maps: [1] normal, [2] specular, [3] gloss
    const float *model;          // MLP parameters (learned capacities)
    float4 *pos4;
    // shading attributes.  sattr null: attr holds three float4 per vertex, every
    // attribute times the vertex's 1/w (u v 1/w n.x | n.y n.z t.x t.y | t.z
    // vd.xyz).  Several jobs of a launch drawing one scene (views): sattr holds
    // that scene's view-independent part once, two float4 per vertex (u v n.x
    // n.y | n.z t.xyz, not times 1/w), and attr one float4 per vertex per job
    // (1/w, vd.xyz times 1/w); shading forms the same products.  The vertex
    // colour (read only where a mesh has no diffuse map) is taken from the
    // mesh's input vertices at shading, times the stored 1/w.
    float4 *attr;
    float4 *sattr;
    uint32_t *trirec;
    uint8_t *alive;
    uint32_t *tile_count;        // exact: counts; single pass: cursors (end = exact counts)
    uint32_t *tile_off;          // num_tiles + 1
    uint32_t *tile_cap;          // single pass capacities
    uint32_t *tile_pool;         // per tile: fallback pool offset, or ~0u (previous-frame policy)
    uint32_t *tile_cursor;       // exact fill cursors / fallback cursors
    const uint32_t *prev_count;  // previous frame's exact per-tile counts
    const float *prev_depth;     // previous frame's per-tile min depth
    uint32_t *queue;
    uint32_t *pool;
    uint2 *ovf;                  // single pass: (tile, id) entries beyond capacity
    uint32_t *survivors;
    uint32_t *big;
    // per screen-list slot: the triangle's depth plane over its pixel range,
    // formed once by the setup pass (screen_plane); read by every resolve tile
    float4 *big_plane;
    FrameCounters *counters;     // num_alive stays 0 here: see alive_part
    FrameCounters *host_counters;  // pinned host copy, written at the end of the launch
    // culling survivors per setup block (launch-wide, indexed by the setup
    // block); the job's blocks are [alive_part0, alive_part0 + alive_parts).
    // Summed when the counters are published: one shared counter updated by
    // every block serializes those updates at its cache line.
    uint32_t *alive_part;
    uint32_t alive_part0, alive_parts;
    float *tile_min_depth;       // per tile min depth written by the resolve pass
    unsigned long long *keys;    // per pixel (z << 32 | ~(id + 1)), reset to depth 1 / none after each resolve
    uint32_t *vis;
    float *depth;
    uint32_t *rgba;
};

}  // namespace rast
