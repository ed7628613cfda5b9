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

// A thread's kPer consecutive words: two 16-byte loads when the buffer is
// 16-byte aligned and the run lies inside it (VEC), else word by word.
template <bool VEC>
__device__ inline void load_run(const uint32_t* __restrict__ in, int64_t base, int64_t n, uint32_t (&v)[kPer]) {
  static_assert(kPer == 8, "two 16-byte loads per thread");
  if (VEC && base + kPer <= n) {
    const uint4 a = *reinterpret_cast<const uint4*>(in + base);
    const uint4 b = *reinterpret_cast<const uint4*>(in + base + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    return;
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) v[j] = base + j < n ? in[base + j] : 0u;
}

template <bool VEC>
__global__ __launch_bounds__(kThreads) void tile_sums(const uint32_t* __restrict__ in, int64_t n,
                                                      uint32_t* __restrict__ sums) {
  __shared__ uint32_t s_wave[kThreads / 64];
  const int64_t base = blockIdx.x * int64_t(kTile) + threadIdx.x * int64_t(kPer);
  uint32_t v[kPer];
  load_run<VEC>(in, base, n, v);
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) s += v[j];
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

template <bool VEC>
__global__ __launch_bounds__(kThreads) void tile_scan(const uint32_t* __restrict__ in, int64_t n,
                                                      const uint32_t* __restrict__ sums,
                                                      uint32_t* __restrict__ out) {
  __shared__ uint32_t s_wave[kThreads / 64];
  const int64_t base = blockIdx.x * int64_t(kTile) + threadIdx.x * int64_t(kPer);
  uint32_t v[kPer];
  load_run<VEC>(in, base, n, v);
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) s += v[j];
  uint32_t run = sums[blockIdx.x] + block_scan(s, s_wave, nullptr);
  uint32_t o[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    o[j] = run;
    run += v[j];
  }
  if (VEC && base + kPer <= n) {
    *reinterpret_cast<uint4*>(out + base) = make_uint4(o[0], o[1], o[2], o[3]);
    *reinterpret_cast<uint4*>(out + base + 4) = make_uint4(o[4], o[5], o[6], o[7]);
    return;
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j)
    if (base + j < n) out[base + j] = o[j];
}

size_t scratch_words(int64_t n) { return size_t(ceil_div(std::max<int64_t>(n, 1), kTile)) + 64; }

void exclusive_u32(hipStream_t st, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* d_total,
                   uint32_t* scratch) {
  const int64_t nt = ceil_div(std::max<int64_t>(n, 1), kTile);
  const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0;
  if (vec) tile_sums<true><<<int(nt), kThreads, 0, st>>>(in, n, scratch);
  else tile_sums<false><<<int(nt), kThreads, 0, st>>>(in, n, scratch);
  scan_sums<<<1, kThreads, 0, st>>>(scratch, nt, d_total);
  if (vec) tile_scan<true><<<int(nt), kThreads, 0, st>>>(in, n, scratch, out);
  else tile_scan<false><<<int(nt), kThreads, 0, st>>>(in, n, scratch, out);
}

}  // namespace scan
