// Stable LSD radix sort of (64-bit key, 32-bit value) pairs for gfx950.
//
// One histogram kernel reads every key once and builds the 256-bin digit
// histograms of all passes.  Each pass is then a single "onesweep" kernel: a
// workgroup takes the next tile (ticket order), ranks its keys stably with
// wave64 ballot matching (8 ballots per batch of 64 keys), publishes its
// per-digit counts, resolves its global offsets with decoupled look-back over
// the preceding tiles, and scatters the tile through LDS so that every digit
// run is written contiguously.
//
// Forward progress: a tile only waits on tiles with smaller tickets, which
// were handed out to already-running workgroups, and each wait is bounded
// (spin limit) -- if the bound expires the kernel sets TRACE_ERR_SPIN in the
// error word and carries on, so the launch always drains.
#pragma once

#include "common.hpp"

namespace onesweep {

constexpr int kThreads = 256;  // 4 waves of 64
constexpr int kWaves = kThreads / 64;
constexpr int kKeysPerThread = 16;
constexpr int kTile = kThreads * kKeysPerThread;  // 4096 keys per tile
constexpr int kRadix = 256;
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;

// Bytes of scratch needed for n keys over `passes` passes (status words and
// tickets; the histogram lives in the same block).
inline size_t status_words(int64_t n, int passes) {
  const int64_t tiles = ceil_div(n, kTile);
  return size_t(passes) * (size_t(tiles) * kRadix + kRadix /*hist*/ + 64 /*tickets*/);
}

// Sort keys[0..n) (values: vals or iota if vals_in == nullptr) on bits
// [0, 8*passes).  When `seg` is given, seg_passes more passes follow whose
// digits are taken from seg[value] (a segment id per original element), so the
// result is ordered by (segment, key) with ties in input order.  Ping-pongs
// between (keys_a, vals_a) and (keys_b, vals_b); returns 0 if the result ends
// in the a-buffers, 1 if in the b-buffers.  `status` must hold
// status_words(n, passes + seg_passes) uint32 and is cleared here.
int sort_pairs(hipStream_t stream, uint64_t* keys_a, uint32_t* vals_a, uint64_t* keys_b,
               uint32_t* vals_b, int64_t n, int passes, uint32_t* status, uint32_t* d_err,
               bool vals_are_iota, const int32_t* seg = nullptr, int seg_passes = 0);

}  // namespace onesweep
