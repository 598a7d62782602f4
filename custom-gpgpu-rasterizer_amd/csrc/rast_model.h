// rast_model.h -- learned bin-capacity model, internal interface used by the
// frame orchestration in rast_api.hip (implementation: rast_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "rast.h"
#include "rast_internal.h"

// Features -> MLP -> per-tile queue capacities for the current frame.
int rast_model_capacities(rast_binmodel *m, const rast::FrameParams &fp, const uint32_t *prev_count,
                          const float *prev_depth, const uint32_t *density, float *features, float margin,
                          uint32_t slack, uint32_t *cap, hipStream_t s);
// Forget a renderer that is being destroyed (null model allowed).
void rast_model_detach_renderer(rast_binmodel *m, rast_renderer *r);
// Capacities from the previous frame's exact counts: ceil(prev * margin) + slack.
hipError_t launch_prevframe_capacity(uint32_t n, const uint32_t *prev, float margin, uint32_t slack, uint32_t *cap,
                                     hipStream_t s);
