// Closest-hit traversal (see traverse.hip).
#pragma once

#include "common.hpp"

namespace traverse {

// Ray record as the C-ABI passes it: origin.xyz, tmin, direction.xyz, tmax.
struct RayIn {
  float4 a;
  float4 b;
};

// Enqueue closest-hit for n rays on st; hit = caller triangle index or -1,
// t = hit distance or tmax.  When counts are non-null, also records per ray
// the inner nodes visited and triangles tested.
int closest_hit(trace_ctx* ctx, const trace_bvh* bvh, const float* d_rays, int64_t n,
                int32_t* d_hit, float* d_t, int32_t* d_nodes, int32_t* d_tris);

// ---- device-side pieces shared with the path tracer -----------------------
struct Ray {
  float ox, oy, oz, tmin;
  float dx, dy, dz, tmax;
};

struct Hit {
  float t;
  int32_t pos;  // sorted triangle position, -1 = miss
  float u, v;   // barycentrics of the accepted hit
  bool front;   // det > 0
  double t64;   // FP64 distance after refine_hit_fp64 (else t)
};

}  // namespace traverse
