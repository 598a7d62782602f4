// rast_launch.h -- host-callable launchers of the kernels in rast_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "rast_internal.h"

namespace rast {

struct FrameBuffers {
    const DevMesh *meshes;
    const uint32_t *vs_block_mesh;
    const uint32_t *tri_block_mesh;
    uint32_t vs_blocks, tri_blocks;
    float4 *pos4;
    float4 *attr;
    uint32_t *trirec;
    uint8_t *alive;
    uint32_t *tile_count;   // exact: counts; single pass: cursors (end = exact counts)
    uint32_t *tile_off;     // num_tiles + 1
    uint32_t *tile_cap;     // single pass capacities (null on the exact path)
    uint32_t *tile_pool;    // per tile: fallback pool offset, or ~0u if the tile did not overflow
    uint32_t *tile_cursor;  // exact fill cursors / fallback cursors
    uint32_t *queue;
    uint32_t queue_cap;
    uint32_t *pool;
    uint32_t pool_cap;
    uint2 *ovf;             // single pass: (tile, id) entries beyond capacity
    uint32_t ovf_cap;
    uint32_t *survivors;
    uint32_t *big;
    uint32_t big_cap;
    FrameCounters *counters;
    uint32_t *density;      // per tile projected-vertex count (learned policy), may be null
    float *tile_min_depth;  // per tile min depth written by the raster, may be null
    unsigned long long *keys;  // per pixel (z << 32 | ~(id + 1)), reset to depth 1 / none after each resolve
    uint32_t *vis;
    float *depth;
    uint32_t *rgba;
};

hipError_t launch_frame_init(const FrameParams &fp, const FrameBuffers &b, bool single_pass, hipStream_t s);
hipError_t launch_clear_keys(unsigned long long *keys, size_t n, hipStream_t s);
hipError_t launch_vertex(const FrameParams &fp, const FrameBuffers &b, hipStream_t s);
hipError_t launch_setup(const FrameParams &fp, const FrameBuffers &b, bool single_pass, hipStream_t s);
hipError_t launch_scan(uint32_t n, const uint32_t *in, uint32_t *out_off, uint32_t *zero_cursor, uint32_t limit,
                       FrameCounters *counters, bool is_capacity, hipStream_t s);
hipError_t launch_fill(const FrameParams &fp, const FrameBuffers &b, hipStream_t s);
hipError_t launch_fallback(const FrameParams &fp, const FrameBuffers &b, hipStream_t s);
hipError_t launch_raster(const FrameParams &fp, const FrameBuffers &b, uint32_t small_limit, hipStream_t s);
hipError_t launch_shade(const FrameParams &fp, const FrameBuffers &b, hipStream_t s);

// Learned bin-capacity model (rast_mlp.hip).
hipError_t launch_tile_features(const FrameParams &fp, const uint32_t *prev_count, const float *prev_depth,
                                const uint32_t *density, float *features, hipStream_t s);
hipError_t launch_mlp_capacity(uint32_t n, const float *features, const float *params, uint32_t hidden,
                               float margin, uint32_t slack, uint32_t *cap, hipStream_t s);

}  // namespace rast
