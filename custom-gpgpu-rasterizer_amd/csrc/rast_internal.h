// rast_internal.h -- device data layout shared by the kernels and the host
// side of librast.so.  gfx950 (CDNA4, wave64) only.
//
// HBM layout of one frame (all device buffers owned by the renderer):
//   ivert    per mesh, 14 floats / vertex (the 56-byte IVertex, as uploaded)
//   pos4     float4 / global vertex: NDC x, y, z and clip w        (16 B)
//   attr     4 x float4 / global vertex: u v 1/w n'.x | n'.yz tan'.xy |
//            tan'.z vd | c.rgb 0 (64 B; PBR shading reads the first 48 B)
//   trirec   16 words / global triangle ID (written for survivors):
//            [0..5] screen x0,y0,x1,y1,x2,y2  [6..8] z0,z1,z2  [9] |area|
//            [10] xmin | ymin << 16  [11] xmax | ymax << 16 (pixel range
//            [min, max)), [12..14] global vertex indices, [15] mesh index
//   alive    u8 / triangle
//   tiles    count / offset / cursor / capacity u32 per tile, queue entries
//            u32 triangle IDs, plus one screen-wide "big" list
//   keys     u64 per pixel: (z bits << 32) | ~(id + 1), min = visible
//            fragment; reset to (1.0, none) by the resolve pass
//   vis, depth  u32 / f32 per pixel, rgba u32 per pixel
#pragma once

#include <cstdint>

namespace rast {

constexpr uint32_t kVisNone = 0xFFFFFFFFu;
constexpr int kIVertexFloats = 14;
constexpr int kTriRecWords = 16;
constexpr int kMaxMeshes = 1 << 16;
constexpr int kNumFeatures = 8;

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
    uint32_t vs_block0;       // first block of this mesh in the vertex launch
    uint32_t tri_block0;      // first block of this mesh in the triangle launch
    int32_t topology;
    int32_t has_diffuse;
    float wvp[16];            // data[col][row]
    float world[16];
    float cam_pos[4];         // camera position (xyz) and forward (xyz) of the frame
    float cam_fwd[4];
    DevTexture tex[4];        // diffuse, normal, specular, gloss (stand-ins filled in)
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
    uint32_t capacity_total;
};

}  // namespace rast
