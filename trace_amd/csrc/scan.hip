// Device-wide exclusive scan (see scan.hpp).
#include "scan.hpp"

namespace scan {

constexpr int kThreads = 256;
constexpr int kPer = 8;
constexpr int kTile = kThreads * kPer;

__device__ inline uint32_t block_scan(uint32_t v, uint32_t* s_wave, uint32_t* total) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[w] = x;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kThreads / 64; ++i) {
    if (i < w) off += s_wave[i];
    tot += s_wave[i];
  }
  __syncthreads();
  if (total) *total = tot;
  return off + x - v;
}

__global__ __launch_bounds__(kThreads) void tile_sums(const uint32_t* __restrict__ in, int64_t n,
                                                      uint32_t* __restrict__ sums) {
  __shared__ uint32_t s_wave[kThreads / 64];
  const int64_t base = blockIdx.x * int64_t(kTile) + threadIdx.x * int64_t(kPer);
  uint32_t s = 0;
  for (int j = 0; j < kPer; ++j)
    if (base + j < n) s += in[base + j];
  uint32_t tot;
  block_scan(s, s_wave, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kThreads) void scan_sums(uint32_t* sums, int64_t nt, uint32_t* d_total) {
  __shared__ uint32_t s_wave[kThreads / 64];
  uint32_t carry = 0;
  for (int64_t b0 = 0; b0 < nt; b0 += kThreads) {
    const int64_t b = b0 + threadIdx.x;
    const uint32_t v = b < nt ? sums[b] : 0u;
    uint32_t tot;
    const uint32_t ex = block_scan(v, s_wave, &tot);
    if (b < nt) sums[b] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && d_total) *d_total = carry;
}

__global__ __launch_bounds__(kThreads) void tile_scan(const uint32_t* __restrict__ in, int64_t n,
                                                      const uint32_t* __restrict__ sums,
                                                      uint32_t* __restrict__ out) {
  __shared__ uint32_t s_wave[kThreads / 64];
  const int64_t base = blockIdx.x * int64_t(kTile) + threadIdx.x * int64_t(kPer);
  uint32_t v[kPer];
  uint32_t s = 0;
  for (int j = 0; j < kPer; ++j) {
    v[j] = base + j < n ? in[base + j] : 0u;
    s += v[j];
  }
  uint32_t run = sums[blockIdx.x] + block_scan(s, s_wave, nullptr);
  for (int j = 0; j < kPer; ++j) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
}

size_t scratch_words(int64_t n) { return size_t(ceil_div(std::max<int64_t>(n, 1), kTile)) + 64; }

void exclusive_u32(hipStream_t st, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* d_total,
                   uint32_t* scratch) {
  const int64_t nt = ceil_div(std::max<int64_t>(n, 1), kTile);
  tile_sums<<<int(nt), kThreads, 0, st>>>(in, n, scratch);
  scan_sums<<<1, kThreads, 0, st>>>(scratch, nt, d_total);
  tile_scan<<<int(nt), kThreads, 0, st>>>(in, n, scratch, out);
}

}  // namespace scan
