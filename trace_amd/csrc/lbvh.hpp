// LBVH build entry (see lbvh.hip).
#pragma once

#include "common.hpp"

namespace lbvh {
// Enqueues the whole build on ctx->stream; returns a TRACE_* status.
int build(trace_ctx* ctx, const trace_mesh* mesh, const double* h_enclosing, int key_mode,
          trace_bvh* bvh);
}  // namespace lbvh
