// rast_model.h -- learned bin-capacity model, internal interface used by the
// frame orchestration in rast_api.hip (implementation: rast_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "rast.h"
#include "rast_internal.h"

// Forget a renderer that is being destroyed (null model allowed).
void rast_model_detach_renderer(rast_binmodel *m, rast_renderer *r);
// Hidden width of a model (16, 32 or 64) and its device parameters.
int rast_model_hidden(const rast_binmodel *m);
const float *rast_model_params(const rast_binmodel *m);
