// rast_launch.h -- host-callable launchers of the frame kernels.
//
// A launch renders one or more frames (jobs) at once: every stage is one
// kernel whose grid is the concatenation of all jobs' blocks for that stage,
// so many small frames fill the chip together.  A single frame is a launch
// of one job (block tables null: every block belongs to job 0).
#pragma once

#include <hip/hip_runtime.h>

#include "rast_internal.h"

namespace rast {

struct Launch {
    const JobDesc *jobs;              // device, njobs entries
    uint32_t njobs;
    const DevMesh *meshes;            // device, all jobs' meshes
    const uint32_t *vs_block_group;   // vertex block -> vertex group
    const uint2 *vs_groups;           // group -> (first index into vs_entries, count)
    const uint32_t *vs_entries;       // the mesh entries of each group (one source mesh, one entry per job)
    const uint32_t *tri_block_mesh;   // triangle block -> mesh entry
    uint32_t vs_blocks, tri_blocks;
    const uint16_t *table[ST_COUNT];  // device block -> job tables (null when njobs == 1)
    uint32_t grid[ST_COUNT];          // blocks of each stage over all jobs (0: stage skipped)
    uint32_t cap_jobs;                // jobs binning in a single pass (capacities placed by the vertex launch)
    uint32_t exact_jobs;              // jobs binning in two passes (count scan)
    int model_hidden;                 // hidden width of the learned jobs' MLP (16, 32 or 64; 0: none)
    uint32_t shade_modes;             // shading instances: bit 0 point, 1 linear; 2, 3 the same for jobs sharing a scene
    uint32_t unshaded_jobs;           // jobs without a shading stage (their counters go out with k_publish)
    // resolve stage: 0 one workgroup per tile; 1 persistent workgroups taking
    // tiles from a counter; 2 the same, shading each tile as it is resolved
    int resolve_mode;
    int fused_shade;                  // 0, or the sampling mode shaded in the resolve pass (1 point, 2 linear)
    uint32_t resident[ST_COUNT];      // persistent resolve: workgroups the GPU holds at once, per stage
    uint32_t *work;                   // device: per-stage work counters, zeroed by the staging upload
};

// Staging bytes in front of the job table: the launch's work counters.
constexpr size_t kStageCounterBytes = 64;

// The staging block of a launch of n jobs: work counters, JobDesc[n], each
// job's FrameCounters (zeros, like the work counters: the upload starts every
// frame's counters at zero), then the mesh table.
static_assert(kStageCounterBytes % 16 == 0 && sizeof(JobDesc) % 16 == 0 && sizeof(DevMesh) % 16 == 0,
              "the staging block is uploaded in 16-byte words");
inline size_t stage_counters_offset(uint32_t n) { return kStageCounterBytes + (size_t)n * sizeof(JobDesc); }
inline size_t stage_meshes_offset(uint32_t n)
{
    return (stage_counters_offset(n) + (size_t)n * sizeof(FrameCounters) + 15) & ~(size_t)15;
}

// Resets a key buffer to depth 1 / no triangle (before a renderer's first frame).
hipError_t launch_clear_keys(unsigned long long *keys, size_t n, hipStream_t s);
hipError_t launch_interleave_maps(const uint32_t *d, const uint32_t *n, const uint32_t *sp, const uint32_t *g,
                                  uint4 *out, size_t count, hipStream_t s);
// with_capacity 0: the vertex pass; 1: the vertex pass with the single-pass
// jobs' capacity workgroups beside it; 2: the capacity workgroups alone.
// Modes 0 and 1 also reset the jobs' per-tile counts and depths (the frame
// counters start at zero with the staging upload).
hipError_t launch_vertex(const Launch &L, int with_capacity, hipStream_t s);
hipError_t launch_scan(const Launch &L, hipStream_t s);  // exact counts
hipError_t launch_setup(const Launch &L, hipStream_t s);
hipError_t launch_fill(const Launch &L, hipStream_t s);
// One tile-size group (ST_RES8 / ST_RES16 / ST_RES32); the groups are independent.
hipError_t launch_resolve(const Launch &L, int stage, hipStream_t s);
// Workgroups of the persistent resolve kernel of a stage the current device
// holds at once (0 on failure).
int resolve_resident_blocks(int stage, int fused_shade);
hipError_t launch_shade(const Launch &L, hipStream_t s);
// Copies a staging block (a multiple of 16 bytes) from pinned host memory.
hipError_t launch_upload(void *dst, const void *src, size_t bytes, hipStream_t s);
hipError_t launch_publish(const Launch &L, hipStream_t s);
// Placement of the previous-frame policy's entries beyond capacity.
hipError_t launch_fallback(const Launch &L, hipStream_t s);
// The DrawTexture / DrawTextureGlobal test hooks: a frame holding one texture
// (vis none, depth 1); mode RAST_DRAW_TEXTURE / RAST_DRAW_TEXTURE_GLOBAL.
hipError_t launch_draw_texture(uint32_t *vis, float *depth, uint32_t *rgba, uint32_t width, uint32_t height,
                               const uint32_t *texels, uint32_t tw, uint32_t th, int mode, bool stretch,
                               int sample_state, hipStream_t s);

// Block -> job lookup of a shared stage grid.  The 16-bit table entry is
// read with a vector load (there is no 16-bit scalar load); it is the same
// for the whole block, and readfirstlane says so, which keeps the job index
// and every job field read through it in scalar registers.
__device__ __forceinline__ uint32_t job_of(const uint16_t *table, const JobDesc *jobs, int stage, uint32_t &local,
                                           uint32_t blk)
{
    if (!table) {
        local = blk;
        return 0;
    }
    const uint32_t j = __builtin_amdgcn_readfirstlane((uint32_t)table[blk]);
    local = blk - jobs[j].block0[stage];
    return j;
}

__device__ __forceinline__ uint32_t job_of(const uint16_t *table, const JobDesc *jobs, int stage, uint32_t &local)
{
    return job_of(table, jobs, stage, local, blockIdx.x);
}

// Training-set features of a rendered frame (rast_mlp.hip).
hipError_t launch_tile_features(const FrameParams &fp, const uint32_t *prev_count, const float *prev_depth,
                                float *features, hipStream_t s);

}  // namespace rast
