// Stable LSD radix sort of (64-bit key, 32-bit value) pairs for gfx950.
//
// One histogram kernel reads every key once and builds the digit histograms
// of all passes.  Each pass is then a single "onesweep" kernel: a workgroup
// takes the next tile (ticket order), ranks its keys stably with wave64 ballot
// matching (one ballot per digit bit per batch of 64 keys), publishes its
// per-digit counts, resolves its global offsets with decoupled look-back over
// the preceding tiles, and scatters the tile through LDS so that every digit
// run is written contiguously.
//
// The digit width (8, 10 or 11 bits) and the keys per thread (4, 8 or 16:
// tiles of 1024 to 4096 keys) are parameters: wider digits mean fewer passes,
// smaller tiles more workgroups for the chip to overlap.  The autotuner picks
// them by timing builds.  Segment digits (forests) are always 8-bit passes.
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
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;
constexpr int kMaxPasses = 12;  // 64-bit keys in 8-bit digits + 3 segment digits, with room

struct Config {
  int bits = 8;  // digit width: 8, 10 or 11
  int kpt = 16;  // keys per thread: 4, 8 or 16
};

bool valid_config(const Config& c);

// Number of key passes for keys of `key_bits` significant bits.
inline int key_passes(int key_bits, const Config& c) { return (key_bits + c.bits - 1) / c.bits; }

// uint32 words of scratch for sorting n keys of key_bits bits (+ seg_passes
// 8-bit segment digits) with configuration c: per-pass tile status words, the
// histograms and the tile tickets.
size_t status_words(int64_t n, int key_bits, int seg_passes, const Config& c);

// Sort keys[0..n) (values: vals or iota if vals_are_iota) on bits
// [0, key_bits).  When `seg` is given, seg_passes more 8-bit passes follow
// whose digits are taken from seg[value] (a segment id per original element),
// so the result is ordered by (segment, key) with ties in input order.
// Ping-pongs between (keys_a, vals_a) and (keys_b, vals_b); returns 0 if the
// result ends in the a-buffers, 1 if in the b-buffers.  `status` must hold
// status_words(n, key_bits, seg_passes, c) uint32 and is cleared here.
int sort_pairs(hipStream_t stream, uint64_t* keys_a, uint32_t* vals_a, uint64_t* keys_b,
               uint32_t* vals_b, int64_t n, int key_bits, uint32_t* status, uint32_t* d_err,
               bool vals_are_iota, const int32_t* seg, int seg_passes, const Config& c);

}  // namespace onesweep
