// Stable LSD onesweep radix sort of (u64 key, u32 value) pairs; see onesweep.hpp.
#include "onesweep.hpp"

namespace onesweep {

// Status words a look-back step loads at once.
constexpr int kLookBatch = 8;

// Words exchanged between workgroups go through agent-scope atomics, which
// bypass the per-XCD L2 on gfx950; workgroup-scope fences only order them.
__device__ inline uint32_t ld_coherent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_coherent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[w] = x;
  __syncthreads();
  uint32_t off = 0;
#pragma unroll
  for (int i = 0; i < kWaves; ++i)
    if (i < w) off += s_wave[i];
  __syncthreads();
  return off + x - v;
}

// What each pass sorts on: `bits` wide digits from bit `shift` of the key, or
// of the element's segment id; its histogram starts at bin `off`.
struct PassDesc {
  int n;
  int shift[kMaxPasses];
  int bits[kMaxPasses];
  int seg[kMaxPasses];
  int off[kMaxPasses];
};

// Digit histograms of all passes in one read of the keys (and segment ids).
__global__ __launch_bounds__(256) void histogram_kernel(const uint64_t* __restrict__ keys, int64_t n,
                                                        const int32_t* __restrict__ seg, PassDesc pd,
                                                        int bins, uint32_t* __restrict__ hist) {
  extern __shared__ uint32_t sh[];  // bins words
  for (int i = threadIdx.x; i < bins; i += blockDim.x) sh[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t k = keys[i];
    const uint32_t s = seg ? uint32_t(seg[i]) : 0u;
    const uint64_t act = __ballot(1);
    const int first = __builtin_amdgcn_readfirstlane(lane);
    for (int p = 0; p < pd.n; ++p) {
      const uint32_t mask = (1u << pd.bits[p]) - 1u;
      const uint32_t d = pd.seg[p] ? (s >> pd.shift[p]) & mask : uint32_t(k >> pd.shift[p]) & mask;
      // spatially coherent keys share their high digits across a wave: one
      // add of the wave's count then replaces 64 adds to one counter
      const uint32_t d_first = __builtin_amdgcn_readfirstlane(d);
      if (__ballot(d == d_first) == act) {
        if (lane == first) atomicAdd(&sh[pd.off[p] + d_first], uint32_t(__popcll(act)));
      } else {
        atomicAdd(&sh[pd.off[p] + d], 1u);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < bins; i += blockDim.x)
    if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

// Digit of an element: from its key, or (segment passes) from seg[value].
__device__ inline uint32_t digit_of(uint64_t key, uint32_t val, int shift, const int32_t* __restrict__ seg,
                                    uint32_t mask) {
  return seg ? (uint32_t(seg[val]) >> shift) & mask : uint32_t(key >> shift) & mask;
}

template <int BITS, int KPT>
__global__ __launch_bounds__(kThreads) void onesweep_pass_kernel(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint64_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, int64_t n, int shift,
    const uint32_t* __restrict__ hist, uint32_t* status, uint32_t* ticket, uint32_t* d_err,
    const int32_t* __restrict__ seg) {
  constexpr int kRadix = 1 << BITS;
  constexpr uint32_t kMask = uint32_t(kRadix - 1);
  constexpr int kTile = kThreads * KPT;
  constexpr int kDpt = kRadix / kThreads;  // digits owned per thread (a contiguous block)
  static_assert(kDpt >= 1 && kDpt <= 8, "1 to 8 digits per thread");
  __shared__ uint64_t s_keys[kTile];
  __shared__ uint32_t s_vals[kTile];
  __shared__ uint32_t s_wh[kWaves * kRadix];  // per-wave digit counters, then scatter bases
  __shared__ uint32_t s_goff[kRadix];         // global position of tile slot 0 per digit
  __shared__ uint32_t s_scan[kWaves];
  __shared__ uint32_t s_tile;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int d0 = tid * kDpt;

  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  for (int i = tid; i < kWaves * kRadix; i += kThreads) s_wh[i] = 0;
  // global start of each digit: exclusive scan of the pass histogram
  uint32_t bucket[kDpt];
  {
    uint32_t hsum = 0;
#pragma unroll
    for (int k = 0; k < kDpt; ++k) {
      bucket[k] = hist[d0 + k];
      hsum += bucket[k];
    }
    uint32_t b = block_exclusive_scan(hsum, s_scan);  // syncs
#pragma unroll
    for (int k = 0; k < kDpt; ++k) {
      const uint32_t c = bucket[k];
      bucket[k] = b;
      b += c;
    }
  }
  const int64_t tile = s_tile;
  const int64_t tile_base = tile * kTile;

  uint64_t key[KPT];
  uint32_t val[KPT];
  uint32_t rank[KPT];
  const uint64_t lt = lanemask_lt();

#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    const int64_t idx = tile_base + w * (KPT * 64) + j * 64 + lane;
    const bool valid = idx < n;
    key[j] = valid ? keys_in[idx] : ~0ull;
    val[j] = valid ? (vals_in ? vals_in[idx] : uint32_t(idx)) : 0u;
  }

  // Stable rank inside the wave: lanes holding the same digit are found with
  // one ballot per digit bit; the highest such lane advances the wave's counter.
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    const int64_t idx = tile_base + w * (KPT * 64) + j * 64 + lane;
    const bool valid = idx < n;
    const uint32_t digit = valid ? digit_of(key[j], val[j], shift, seg, kMask) : kMask;
    uint64_t m = ballot64(valid);
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
      const bool bit = (digit >> b) & 1u;
      const uint64_t bb = ballot64(bit);
      m &= bit ? bb : ~bb;
    }
    const uint32_t before = __popcll(m & lt);
    const uint32_t total = __popcll(m);
    const uint32_t base = s_wh[w * kRadix + digit];
    const bool leader = valid && (m >> lane) == 1ull;
    if (leader) s_wh[w * kRadix + digit] = base + total;
    rank[j] = base + before;
  }
  __syncthreads();

  // Thread t owns digits [t * kDpt, (t + 1) * kDpt): wave offsets, tile-local
  // starts (folded into the per-wave scatter bases), look-back.
  uint32_t tot[kDpt];
  uint32_t tsum = 0;
#pragma unroll
  for (int k = 0; k < kDpt; ++k) {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
      const uint32_t c = s_wh[i * kRadix + d0 + k];
      s_wh[i * kRadix + d0 + k] = acc;
      acc += c;
    }
    tot[k] = acc;
    tsum += acc;
  }
  uint32_t local[kDpt];
  {
    uint32_t b = block_exclusive_scan(tsum, s_scan);  // syncs
#pragma unroll
    for (int k = 0; k < kDpt; ++k) {
      local[k] = b;
      b += tot[k];
#pragma unroll
      for (int i = 0; i < kWaves; ++i) s_wh[i * kRadix + d0 + k] += local[k];
    }
  }

  uint32_t excl[kDpt];
  uint32_t* mine = status + size_t(tile) * kRadix + d0;
  if (tile == 0) {
#pragma unroll
    for (int k = 0; k < kDpt; ++k) {
      excl[k] = 0;
      st_coherent(mine + k, kFlagPrefix | tot[k]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < kDpt; ++k) st_coherent(mine + k, kFlagAgg | tot[k]);
    // every owned digit walks back over the preceding tiles, kLookBatch
    // status words per step: the batch's loads are in flight together, and
    // its words are then taken in tile order up to the first inclusive prefix
    // (done) or the first tile that has published nothing yet (wait there)
    int64_t at[kDpt];
    uint32_t pending = (1u << kDpt) - 1u;
#pragma unroll
    for (int k = 0; k < kDpt; ++k) {
      at[k] = tile - 1;
      excl[k] = 0;
    }
    uint32_t spins = 0;
    while (pending) {
      bool waited = false;
#pragma unroll
      for (int k = 0; k < kDpt; ++k) {
        if (!((pending >> k) & 1u)) continue;
        uint32_t v[kLookBatch];
#pragma unroll
        for (int b = 0; b < kLookBatch; ++b)
          v[b] = at[k] - b >= 0 ? ld_coherent(status + size_t(at[k] - b) * kRadix + d0 + k) : kFlagPrefix;
        bool stop = false;
#pragma unroll
        for (int b = 0; b < kLookBatch; ++b) {
          if (stop) continue;
          const uint32_t f = v[b] & ~kValueMask;
          if (f == 0) {  // tile at[k] has published nothing yet
            waited = true;
            stop = true;
            continue;
          }
          excl[k] += v[b] & kValueMask;
          if (f == kFlagPrefix || at[k] == 0) {
            pending &= ~(1u << k);
            stop = true;
          } else {
            --at[k];
          }
        }
      }
      if (waited && pending) {
        if (++spins > kSpinLimit) {
          atomicOr(d_err, TRACE_ERR_SPIN);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
#pragma unroll
    for (int k = 0; k < kDpt; ++k) st_coherent(mine + k, kFlagPrefix | ((excl[k] + tot[k]) & kValueMask));
  }
#pragma unroll
  for (int k = 0; k < kDpt; ++k) s_goff[d0 + k] = bucket[k] + excl[k] - local[k];
  __syncthreads();

  // Scatter into LDS in tile-sorted order.
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    const int64_t idx = tile_base + w * (KPT * 64) + j * 64 + lane;
    if (idx < n) {
      const uint32_t digit = digit_of(key[j], val[j], shift, seg, kMask);
      const uint32_t pos = s_wh[w * kRadix + digit] + rank[j];
      s_keys[pos] = key[j];
      s_vals[pos] = val[j];
    }
  }
  __syncthreads();

  const int64_t remaining = n - tile_base;
  const int count = remaining < kTile ? int(remaining) : kTile;
  for (int i = tid; i < count; i += kThreads) {
    const uint64_t k = s_keys[i];
    const uint32_t digit = digit_of(k, s_vals[i], shift, seg, kMask);
    const uint32_t g = s_goff[digit] + uint32_t(i);
    if (g < uint64_t(n)) {  // always true unless a look-back wait expired
      keys_out[g] = k;
      vals_out[g] = s_vals[i];
    }
  }
}

bool valid_config(const Config& c) {
  const bool bits_ok = c.bits == 8 || c.bits == 10 || c.bits == 11;
  const bool kpt_ok = c.kpt == 4 || c.kpt == 8 || c.kpt == 16;
  // 11-bit digits keep 2048 counters per wave in LDS: tiles of 1024 keys only
  return bits_ok && kpt_ok && !(c.bits == 11 && c.kpt != 4) && !(c.bits == 10 && c.kpt == 16);
}

namespace {

// Pass p of a sort with k key passes: its digit width, shift and source.
struct Pass {
  int bits, shift;
  bool seg;
};
Pass pass_of(int p, int kp, const Config& c) {
  if (p < kp) return Pass{c.bits, c.bits * p, false};
  return Pass{8, 8 * (p - kp), true};
}

template <int BITS, int KPT>
void launch(int64_t tiles, hipStream_t st, const uint64_t* kin, const uint32_t* vin, uint64_t* kout,
            uint32_t* vout, int64_t n, int shift, const uint32_t* hist, uint32_t* status, uint32_t* ticket,
            uint32_t* d_err, const int32_t* seg) {
  onesweep_pass_kernel<BITS, KPT><<<int(tiles), kThreads, 0, st>>>(kin, vin, kout, vout, n, shift, hist, status,
                                                                    ticket, d_err, seg);
}

}  // namespace

size_t status_words(int64_t n, int key_bits, int seg_passes, const Config& c) {
  const int kp = key_passes(key_bits, c);
  const int64_t tiles = ceil_div(std::max<int64_t>(n, 1), int64_t(kThreads) * c.kpt);
  size_t words = 64;  // tickets
  for (int p = 0; p < kp + seg_passes; ++p) {
    const size_t radix = size_t(1) << pass_of(p, kp, c).bits;
    words += size_t(tiles) * radix + radix;  // tile status + histogram
  }
  return words;
}

int sort_pairs(hipStream_t stream, uint64_t* keys_a, uint32_t* vals_a, uint64_t* keys_b,
               uint32_t* vals_b, int64_t n, int key_bits, uint32_t* status, uint32_t* d_err,
               bool vals_are_iota, const int32_t* seg, int seg_passes, const Config& c) {
  if (n <= 0) return 0;
  if (!seg) seg_passes = 0;
  const int kp = key_passes(key_bits, c);
  const int all = kp + seg_passes;
  const int64_t tiles = ceil_div(n, int64_t(kThreads) * c.kpt);
  PassDesc pd{};
  pd.n = all;
  size_t status_off[kMaxPasses];
  size_t so = 0;
  int bins = 0;
  for (int p = 0; p < all; ++p) {
    const Pass ps = pass_of(p, kp, c);
    pd.shift[p] = ps.shift;
    pd.bits[p] = ps.bits;
    pd.seg[p] = ps.seg ? 1 : 0;
    pd.off[p] = bins;
    bins += 1 << ps.bits;
    status_off[p] = so;
    so += size_t(tiles) << ps.bits;
  }
  uint32_t* hist = status + so;
  uint32_t* tickets = hist + bins;
  fill_async<uint32_t>(stream, status, status_words(n, key_bits, seg_passes, c), 0u);
  {
    const int blocks = int(std::min<int64_t>(ceil_div(n, 256 * 8), 1024));
    histogram_kernel<<<blocks, 256, size_t(bins) * sizeof(uint32_t), stream>>>(keys_a, n, seg_passes ? seg : nullptr,
                                                                                pd, bins, hist);
  }
  uint64_t* kin = keys_a;
  uint32_t* vin = vals_are_iota ? nullptr : vals_a;
  uint64_t* kout = keys_b;
  uint32_t* vout = vals_b;
  int where = 0;
  for (int p = 0; p < all; ++p) {
    const Pass ps = pass_of(p, kp, c);
    const uint32_t* h = hist + pd.off[p];
    uint32_t* s = status + status_off[p];
    const int32_t* sg = ps.seg ? seg : nullptr;
#define TRACE_SORT_LAUNCH(B, K) \
  launch<B, K>(tiles, stream, kin, vin, kout, vout, n, ps.shift, h, s, tickets + p, d_err, sg)
    if (ps.bits == 8 && c.kpt == 16) TRACE_SORT_LAUNCH(8, 16);
    else if (ps.bits == 8 && c.kpt == 8) TRACE_SORT_LAUNCH(8, 8);
    else if (ps.bits == 8) TRACE_SORT_LAUNCH(8, 4);
    else if (ps.bits == 10 && c.kpt == 8) TRACE_SORT_LAUNCH(10, 8);
    else if (ps.bits == 10) TRACE_SORT_LAUNCH(10, 4);
    else TRACE_SORT_LAUNCH(11, 4);
#undef TRACE_SORT_LAUNCH
    where ^= 1;
    if (where) {
      kin = keys_b; vin = vals_b; kout = keys_a; vout = vals_a;
    } else {
      kin = keys_a; vin = vals_a; kout = keys_b; vout = vals_b;
    }
  }
  return where;
}

}  // namespace onesweep
