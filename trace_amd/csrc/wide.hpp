// 4-wide traversal tree derived from a binary tree (see wide.hip).
#pragma once

#include "common.hpp"

namespace wide {
// Enqueue the widening of bvh's binary tree (LBVH, forest or PHR layout) into
// bvh->wnodes and set wroot in its header(s).  The binary inner-node count is
// read on the device from the header; bvh->nodes' allocation bounds it.
// links (optional): the binary tree's parent links as the LBVH build leaves
// them, (parent << 1) | slot per inner node and -1 at roots; without them the
// links are derived from the nodes.  leaf_links (optional, with links): the
// LBVH leaves' links in the same form, one per distinct key, whose triangle
// runs are bvh->run_start; they let the split-leaf sizes come from the runs
// instead of a read of every node.
// ws: as lbvh::build's (a fixed workspace allocates nothing).
int widen(trace_ctx* ctx, trace_bvh* bvh, const int32_t* links = nullptr, const Workspace* ws = nullptr,
          const int32_t* leaf_links = nullptr);
// Scratch bytes widen() takes for a binary tree of `inner_cap` inner-node
// slots, `n_roots` trees and `n_prims` triangles, and the wide nodes it writes.
size_t scratch_bytes(int64_t inner_cap, int32_t n_roots);
int64_t wide_nodes(int64_t inner_cap, int64_t n_prims);
}  // namespace wide
