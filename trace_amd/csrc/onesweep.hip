// Stable LSD onesweep radix sort of (u64 key, u32 value) pairs; see onesweep.hpp.
#include "onesweep.hpp"

namespace onesweep {

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

// Digit histograms of all passes in one read of the keys (and segment ids).
__global__ __launch_bounds__(256) void histogram_kernel(const uint64_t* __restrict__ keys,
                                                        int64_t n, int passes,
                                                        const int32_t* __restrict__ seg,
                                                        int seg_passes,
                                                        uint32_t* __restrict__ hist) {
  __shared__ uint32_t sh[11 * kRadix];  // 8 key digits + up to 3 segment digits
  const int all = passes + seg_passes;
  for (int i = threadIdx.x; i < all * kRadix; i += blockDim.x) sh[i] = 0;
  __syncthreads();
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t k = keys[i];
    for (int p = 0; p < passes; ++p) atomicAdd(&sh[p * kRadix + ((k >> (8 * p)) & 255)], 1u);
    if (seg_passes) {
      const uint32_t s = uint32_t(seg[i]);
      for (int p = 0; p < seg_passes; ++p)
        atomicAdd(&sh[(passes + p) * kRadix + ((s >> (8 * p)) & 255)], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < all * kRadix; i += blockDim.x)
    if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

// Digit of an element: from its key, or (segment passes) from seg[value].
__device__ inline uint32_t digit_of(uint64_t key, uint32_t val, int shift,
                                    const int32_t* __restrict__ seg) {
  return seg ? (uint32_t(seg[val]) >> shift) & 255u : uint32_t(key >> shift) & 255u;
}

__global__ __launch_bounds__(kThreads) void onesweep_pass_kernel(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint64_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, int64_t n, int shift,
    const uint32_t* __restrict__ hist, uint32_t* status, uint32_t* ticket, uint32_t* d_err,
    const int32_t* __restrict__ seg) {
  __shared__ uint64_t s_keys[kTile];
  __shared__ uint32_t s_vals[kTile];
  __shared__ uint32_t s_wh[kWaves * kRadix];  // per-wave digit counters, then offsets
  __shared__ uint32_t s_bucket[kRadix];       // global start of each digit
  __shared__ uint32_t s_local[kRadix];        // start of each digit inside the tile
  __shared__ uint32_t s_goff[kRadix];         // global position of local slot 0 per digit
  __shared__ uint32_t s_scan[kWaves];
  __shared__ uint32_t s_tile;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;

  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  for (int i = tid; i < kWaves * kRadix; i += kThreads) s_wh[i] = 0;
  const uint32_t bucket = block_exclusive_scan(hist[tid], s_scan);  // syncs
  s_bucket[tid] = bucket;
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t tile_base = tile * kTile;

  uint64_t key[kKeysPerThread];
  uint32_t val[kKeysPerThread];
  uint32_t rank[kKeysPerThread];
  const uint64_t lt = lanemask_lt();

#pragma unroll
  for (int j = 0; j < kKeysPerThread; ++j) {
    const int64_t idx = tile_base + w * (kKeysPerThread * 64) + j * 64 + lane;
    const bool valid = idx < n;
    key[j] = valid ? keys_in[idx] : ~0ull;
    val[j] = valid ? (vals_in ? vals_in[idx] : uint32_t(idx)) : 0u;
  }

  // Stable rank inside the wave: lanes holding the same digit are found with
  // 8 ballots; the highest such lane advances the wave's counter for it.
#pragma unroll
  for (int j = 0; j < kKeysPerThread; ++j) {
    const int64_t idx = tile_base + w * (kKeysPerThread * 64) + j * 64 + lane;
    const bool valid = idx < n;
    const uint32_t digit = valid ? digit_of(key[j], val[j], shift, seg) : 255u;
    uint64_t m = ballot64(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
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

  // Thread d owns digit d: wave offsets, tile-local start, look-back.
  {
    const int d = tid;
    uint32_t c[kWaves];
    uint32_t total = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
      c[i] = s_wh[i * kRadix + d];
      total += c[i];
    }
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
      s_wh[i * kRadix + d] = acc;
      acc += c[i];
    }
    const uint32_t local = block_exclusive_scan(total, s_scan);  // syncs
    s_local[d] = local;

    uint32_t excl = 0;
    uint32_t* mine = status + size_t(tile) * kRadix + d;
    if (tile == 0) {
      st_coherent(mine, kFlagPrefix | total);
    } else {
      st_coherent(mine, kFlagAgg | total);
      int64_t t = tile - 1;
      uint32_t spins = 0;
      while (true) {
        const uint32_t v = ld_coherent(status + size_t(t) * kRadix + d);
        const uint32_t f = v & ~kValueMask;
        if (f == 0) {
          if (++spins > kSpinLimit) {
            atomicOr(d_err, TRACE_ERR_SPIN);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        excl += v & kValueMask;
        if (f == kFlagPrefix || t == 0) break;
        --t;
      }
      st_coherent(mine, kFlagPrefix | ((excl + total) & kValueMask));
    }
    s_goff[d] = s_bucket[d] + excl - local;
  }
  __syncthreads();

  // Scatter into LDS in tile-sorted order.
#pragma unroll
  for (int j = 0; j < kKeysPerThread; ++j) {
    const int64_t idx = tile_base + w * (kKeysPerThread * 64) + j * 64 + lane;
    if (idx < n) {
      const uint32_t digit = digit_of(key[j], val[j], shift, seg);
      const uint32_t pos = s_local[digit] + s_wh[w * kRadix + digit] + rank[j];
      s_keys[pos] = key[j];
      s_vals[pos] = val[j];
    }
  }
  __syncthreads();

  const int64_t remaining = n - tile_base;
  const int count = remaining < kTile ? int(remaining) : kTile;
  for (int i = tid; i < count; i += kThreads) {
    const uint64_t k = s_keys[i];
    const uint32_t digit = digit_of(k, s_vals[i], shift, seg);
    const uint32_t g = s_goff[digit] + uint32_t(i);
    if (g < uint64_t(n)) {  // always true unless a look-back wait expired
      keys_out[g] = k;
      vals_out[g] = s_vals[i];
    }
  }
}

int sort_pairs(hipStream_t stream, uint64_t* keys_a, uint32_t* vals_a, uint64_t* keys_b,
               uint32_t* vals_b, int64_t n, int passes, uint32_t* status, uint32_t* d_err,
               bool vals_are_iota, const int32_t* seg, int seg_passes) {
  if (n <= 0) return 0;
  if (!seg) seg_passes = 0;
  const int all = passes + seg_passes;
  const int64_t tiles = ceil_div(n, kTile);
  const size_t per_pass = size_t(tiles) * kRadix;
  uint32_t* hist = status + size_t(all) * per_pass;
  uint32_t* tickets = hist + size_t(all) * kRadix;
  const size_t words = status_words(n, all);
  fill_async<uint32_t>(stream, status, words, 0u);
  {
    const int blocks = int(std::min<int64_t>(ceil_div(n, 256 * 8), 1024));
    histogram_kernel<<<blocks, 256, 0, stream>>>(keys_a, n, passes, seg, seg_passes, hist);
  }
  uint64_t* kin = keys_a;
  uint32_t* vin = vals_are_iota ? nullptr : vals_a;
  uint64_t* kout = keys_b;
  uint32_t* vout = vals_b;
  int where = 0;
  for (int p = 0; p < all; ++p) {
    const bool seg_pass = p >= passes;
    onesweep_pass_kernel<<<int(tiles), kThreads, 0, stream>>>(
        kin, vin, kout, vout, n, 8 * (seg_pass ? p - passes : p), hist + size_t(p) * kRadix,
        status + size_t(p) * per_pass, tickets + p, d_err, seg_pass ? seg : nullptr);
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
