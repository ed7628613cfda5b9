// Progressive Hierarchical Refinement on the GPU (see phr.hip).
#pragma once

#include "common.hpp"

namespace phr {

// One node of the PHR tree as the reference builds it (n-ary, here at most 2
// children): a branch whose box is its cut's bounding, or a leaf made from a
// list of LBVH nodes (phr.go:340-352).  Triangles of the subtree occupy
// [prim_off, prim_off + prim_cnt) of the PHR triangle order.
struct OutNode {
  int32_t kind;      // 0 branch, 1 leaf
  int32_t nchild;    // branch: 1 or 2
  int32_t child[2];  // branch: output node indices
  int32_t list_off;  // leaf: LBVH node ids in list[list_off, list_off + list_cnt)
  int32_t list_cnt;
  int32_t prim_off;
  int32_t prim_cnt;
  int32_t level;     // depth below the root
  float box[6];
  int32_t pad;
};
static_assert(sizeof(OutNode) == 64, "OutNode is 64 bytes");

// How build() runs:
//  kFirstBuild   -- a new tree, sized to what it holds (synchronous);
//  kRebuild      -- into an existing tree (synchronous); a binary rebuild sizes
//                   its scratch for any frame of the same triangle count, its
//                   node array for an eighth more branches than it makes, and
//                   keeps the threshold powers on the device;
//  kGraphRebuild -- a binary rebuild while the context stream is captured: no
//                   host reads, every launch sized by capacities, the job
//                   levels of the last eager rebuild plus kGraphSpareLevels
//                   recorded; a replay whose frame needs more levels or
//                   branches, or makes a deeper tree, or has an initial cut of
//                   one node, sets TRACE_ERR_BOUND.
enum Mode { kFirstBuild = 0, kRebuild = 1, kGraphRebuild = 2 };
constexpr int kGraphSpareLevels = 4;

// BuildFromAuxilary(aux) with PhrBuilder{alpha, delta, branching} and the
// initial-cut cap max_cut (phr.go:11, 47-84).
int build(trace_ctx* ctx, const trace_bvh* aux, double alpha, double delta, int branching,
          int32_t max_cut, trace_bvh* out, Mode mode);

}  // namespace phr
