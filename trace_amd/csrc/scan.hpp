// Device-wide exclusive scan of uint32 values (reduce-then-scan, three
// kernels, no host round trip).  The total lands in *d_total.
#pragma once

#include "common.hpp"

namespace scan {
// Scratch words needed for n values.
size_t scratch_words(int64_t n);
void exclusive_u32(hipStream_t st, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* d_total,
                   uint32_t* scratch);
}  // namespace scan
