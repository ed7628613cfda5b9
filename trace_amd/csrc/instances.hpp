// Two-level trees: a top level over instances of bottom-level trees (see instances.hip).
#pragma once

#include "common.hpp"

namespace instances {
// Closest hit of n rays on an instanced tree (bvh->instanced): hit = flattened
// triangle index or -1, t = distance or tmax.
int closest_hit(trace_ctx* ctx, const trace_bvh* bvh, const float* d_rays, int64_t n, int32_t* d_hit,
                float* d_t);
}  // namespace instances
