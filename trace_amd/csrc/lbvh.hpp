// LBVH build entry (see lbvh.hip).
#pragma once

#include "common.hpp"

namespace lbvh {
// Enqueues the whole build on ctx->stream; returns a TRACE_* status.  With a
// fixed workspace (ws) the build allocates nothing: its scratch comes from
// ws and the tree's arrays must already be sized for the mesh.
int build(trace_ctx* ctx, const trace_mesh* mesh, const double* h_enclosing, int key_mode,
          trace_bvh* bvh, const Workspace* ws = nullptr);
// Scratch bytes build() takes for this mesh, key mode and the context's sort
// configuration (the widening pass that follows reuses the same memory:
// wide::scratch_bytes).
size_t scratch_bytes(const trace_ctx* ctx, const trace_mesh* mesh, int key_mode);
}  // namespace lbvh
