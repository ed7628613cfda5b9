// Morton LBVH build for gfx950.
//
// Pipeline (all on the context stream, no host round trip):
//   enclosing_kernel  FP64 scene box = union of triangle boxes       bounding.go:19-25
//   header_kernel     root box, quantisation box                      lbvh.go:13-15
//   morton_kernel     63-bit FP64 Morton code per triangle             lbvh.go:64-73, morton.go:6-20
//   onesweep sort     stable radix sort of (key, triangle index)       lbvh.go:216-301 (bucket sort)
//   unique kernels    equal-key runs -> one leaf each                  lbvh.go:134-145, 179-181
//   karras_kernel     binary radix tree over the unique keys           lbvh.go:147-203 (top-down split)
//   reorder_kernel    FP32 triangles and boxes in sorted order
//   refit_kernel      bottom-up AABBs with per-node arrival counters   bvh.go:104-121, 201-222
//
// The reference builds the tree top-down: a key range whose first and last
// codes are equal becomes one leaf; otherwise the range splits at the first
// key that has the highest differing bit set.  That is exactly the binary
// radix tree of Karras (2012) over the distinct keys, which every inner node
// can find independently; the leaf of distinct key u is the run of triangles
// sharing it.
#include <algorithm>

#include "lbvh.hpp"
#include "onesweep.hpp"
#include "wide.hpp"

namespace lbvh {

__device__ inline uint64_t ord_of(double x) {
  const uint64_t b = __double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double double_of(uint64_t o) {
  const uint64_t b = (o >> 63) ? (o & 0x7fffffffffffffffull) : ~o;
  return __longlong_as_double(b);
}

constexpr int kEncBlocks = 2048;

// Stage 1 of the scene box: each workgroup reduces a grid-stride share of the
// triangles (min/max are exact, so any order gives the reference's box) and
// writes one partial box as order-preserving u64 words; no atomics.
// A tile of 256 triangles' positions (256 x 72 bytes) staged through LDS with
// 16-byte loads: each wave reads whole lines instead of nine 8-byte loads per
// lane at a 72-byte stride.  Returns the tile's triangle count.
constexpr int kPosTile = 256;
__device__ inline int stage_positions(const double* __restrict__ pos, int64_t n, int64_t base, double* s_pos) {
  const int64_t left = n - base;
  const int cnt = left < kPosTile ? int(left) : kPosTile;
  const int words = 9 * cnt;  // doubles; base is even, so the tile starts on a 16-byte boundary
  const double2* src = reinterpret_cast<const double2*>(pos + 9 * base);
  double2* dst = reinterpret_cast<double2*>(s_pos);
  for (int k = threadIdx.x; k < words / 2; k += blockDim.x) dst[k] = src[k];
  if ((words & 1) && threadIdx.x == 0) s_pos[words - 1] = pos[9 * base + words - 1];
  __syncthreads();
  return cnt;
}

__global__ __launch_bounds__(256) void enclosing_kernel(const double* __restrict__ pos, int64_t n,
                                                        unsigned long long* __restrict__ partial) {
  __shared__ uint64_t s_part[6][4];
  __shared__ double s_pos[9 * kPosTile];
  uint64_t lo[3] = {~0ull, ~0ull, ~0ull};
  uint64_t hi[3] = {0ull, 0ull, 0ull};
  for (int64_t base = blockIdx.x * int64_t(kPosTile); base < n; base += int64_t(gridDim.x) * kPosTile) {
    const int cnt = stage_positions(pos, n, base, s_pos);
    if (int(threadIdx.x) < cnt) {
      const double* p = s_pos + 9 * threadIdx.x;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double l, h;
        prim_bounds(p, a, l, h);
        const uint64_t mn = ord_of(l);
        const uint64_t mx = ord_of(h);
        lo[a] = mn < lo[a] ? mn : lo[a];
        hi[a] = mx > hi[a] ? mx : hi[a];
      }
    }
    __syncthreads();  // the next tile reuses s_pos
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t l = __shfl_xor(lo[a], o, 64);
      const uint64_t h = __shfl_xor(hi[a], o, 64);
      lo[a] = l < lo[a] ? l : lo[a];
      hi[a] = h > hi[a] ? h : hi[a];
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      s_part[a][w] = lo[a];
      s_part[3 + a][w] = hi[a];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    uint64_t v = s_part[a][0];
    for (int k = 1; k < 4; ++k) v = a < 3 ? (s_part[a][k] < v ? s_part[a][k] : v)
                                          : (s_part[a][k] > v ? s_part[a][k] : v);
    partial[6 * blockIdx.x + a] = v;
  }
}

// Stage 2: reduces the partial boxes, writes the scene box into the header
// and, unless the caller supplied one, the quantisation box used by the
// Morton kernel.
__global__ void header_kernel(const unsigned long long* __restrict__ partial, int nparts,
                              BvhHeader* hdr, double* qbox, int caller_box, int32_t n) {
  __shared__ uint64_t s_box[6][4];
  uint64_t v[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) v[a] = a < 3 ? ~0ull : 0ull;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      const uint64_t p = partial[6 * k + a];
      v[a] = a < 3 ? (p < v[a] ? p : v[a]) : (p > v[a] ? p : v[a]);
    }
  }
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t p = __shfl_xor(v[a], o, 64);
      v[a] = a < 3 ? (p < v[a] ? p : v[a]) : (p > v[a] ? p : v[a]);
    }
    if ((threadIdx.x & 63) == 0) s_box[a][threadIdx.x >> 6] = v[a];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int a = 0; a < 6; ++a) {
    uint64_t r = s_box[a][0];
    for (int k = 1; k < int(blockDim.x >> 6); ++k)
      r = a < 3 ? (s_box[a][k] < r ? s_box[a][k] : r) : (s_box[a][k] > r ? s_box[a][k] : r);
    s_box[a][0] = r;
  }
  hdr->n_prims = n;
  for (int a = 0; a < 3; ++a) {
    const double lo = double_of(s_box[a][0]);
    const double hi = double_of(s_box[3 + a][0]);
    hdr->enc_lo[a] = lo;
    hdr->enc_hi[a] = hi;
    hdr->root_lo[a] = f32_down(lo);
    hdr->root_hi[a] = f32_up(hi);
    if (!caller_box) {
      qbox[a] = lo;
      qbox[3 + a] = hi;
    }
  }
}

// ---- forests: one scene box per segment ---------------------------------------
__global__ void seg_enc_init_kernel(unsigned long long* enc, int32_t nseg) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 6 * nseg) enc[i] = (i % 6) < 3 ? ~0ull : 0ull;
}

constexpr int kSegTile = 4096;

// Each workgroup covers kSegTile consecutive triangles.  Segments are
// contiguous, so a tile lies in one segment unless it straddles a boundary:
// such tiles (at most nseg - 1 of them) fall back to per-triangle atomics.
__global__ __launch_bounds__(256) void seg_enclosing_kernel(const double* __restrict__ pos,
                                                            int64_t n,
                                                            const int32_t* __restrict__ tseg,
                                                            unsigned long long* enc) {
  __shared__ uint64_t s_part[6][4];
  const int64_t t0 = blockIdx.x * int64_t(kSegTile);
  const int64_t t1 = t0 + kSegTile < n ? t0 + kSegTile : n;
  const int32_t s0 = tseg[t0];
  const bool uniform = s0 == tseg[t1 - 1];
  uint64_t lo[3] = {~0ull, ~0ull, ~0ull};
  uint64_t hi[3] = {0ull, 0ull, 0ull};
  for (int64_t i = t0 + threadIdx.x; i < t1; i += blockDim.x) {
    const double* p = pos + 9 * i;
    for (int a = 0; a < 3; ++a) {
      double l, h;
      prim_bounds(p, a, l, h);
      const uint64_t mn = ord_of(l);
      const uint64_t mx = ord_of(h);
      if (uniform) {
        lo[a] = mn < lo[a] ? mn : lo[a];
        hi[a] = mx > hi[a] ? mx : hi[a];
      } else {
        atomicMin(&enc[6 * size_t(tseg[i]) + a], (unsigned long long)mn);
        atomicMax(&enc[6 * size_t(tseg[i]) + 3 + a], (unsigned long long)mx);
      }
    }
  }
  if (!uniform) return;  // uniform is the same for the whole workgroup
  for (int a = 0; a < 3; ++a) {
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t l = __shfl_xor(lo[a], o, 64);
      const uint64_t h = __shfl_xor(hi[a], o, 64);
      lo[a] = l < lo[a] ? l : lo[a];
      hi[a] = h > hi[a] ? h : hi[a];
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    for (int a = 0; a < 3; ++a) {
      s_part[a][w] = lo[a];
      s_part[3 + a][w] = hi[a];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    uint64_t v = s_part[a][0];
    for (int k = 1; k < 4; ++k)
      v = a < 3 ? (s_part[a][k] < v ? s_part[a][k] : v) : (s_part[a][k] > v ? s_part[a][k] : v);
    if (a < 3) atomicMin(&enc[6 * size_t(s0) + a], (unsigned long long)v);
    else atomicMax(&enc[6 * size_t(s0) + a], (unsigned long long)v);
  }
}

// Per segment: FP64 scene box, outward FP32 root box, quantisation box.
__global__ void seg_header_kernel(const unsigned long long* __restrict__ enc, int32_t nseg,
                                  BvhHeader* __restrict__ seg_hdr, double* __restrict__ qbox,
                                  BvhHeader* __restrict__ hdr, int32_t n) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) hdr->n_prims = n;
  if (s >= nseg) return;
  BvhHeader& h = seg_hdr[s];
  for (int a = 0; a < 3; ++a) {
    const double lo = double_of(enc[6 * size_t(s) + a]);
    const double hi = double_of(enc[6 * size_t(s) + 3 + a]);
    h.enc_lo[a] = lo;
    h.enc_hi[a] = hi;
    h.root_lo[a] = f32_down(lo);
    h.root_hi[a] = f32_up(hi);
    qbox[6 * size_t(s) + a] = lo;
    qbox[6 * size_t(s) + 3 + a] = hi;
  }
  h.n_prims = n;
  h.depth = 0;
}

__device__ inline uint64_t split_by3(uint64_t a) {
  // Spread the low 21 bits of a so that bit k lands on bit 3k.
  uint64_t x = a & 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

// q = uint64(|enc.min - c| / (extent / (2^21 - 1))), evaluated in FP64 with
// the reference's operation order (lbvh.go:64-73).  No contraction: every
// operation is rounded on its own, as Go does on amd64.  A zero extent gives
// 0/0; the reference's result is platform dependent and here q = 0.
__device__ inline uint64_t quantise(double enc_lo, double extent, double c) {
#pragma clang fp contract(off)
  const double delta = fabs(enc_lo - c);
  const double step = extent / 2097151.0;
  const double f = delta / step;
  if (!(f >= 0.0)) return 0;      // NaN
  if (f >= 18446744073709551615.0) return ~0ull;
  return (uint64_t)f;
}

// tseg (segmented meshes): each triangle quantises in its own segment's box.
__global__ __launch_bounds__(kPosTile) void morton_kernel(const double* __restrict__ pos, int64_t n,
                                                          const double* __restrict__ qbox_all,
                                                          const int32_t* __restrict__ tseg, int key_mode,
                                                          uint64_t* __restrict__ keys) {
#pragma clang fp contract(off)
  __shared__ double s_pos[9 * kPosTile];
  const int64_t base = blockIdx.x * int64_t(kPosTile);
  stage_positions(pos, n, base, s_pos);
  const int64_t i = base + threadIdx.x;
  if (i >= n) return;
  const double* qbox = tseg ? qbox_all + 6 * size_t(tseg[i]) : qbox_all;
  const double* p = s_pos + 9 * threadIdx.x;
  uint64_t q[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double lo, hi;
    prim_bounds(p, a, lo, hi);
    const double c = (lo + hi) * 0.5;  // aabb.update, bounding.go:36-43
    q[a] = quantise(qbox[a], qbox[3 + a] - qbox[a], c);
  }
  uint64_t code = split_by3(q[0]) | (split_by3(q[1]) << 1) | (split_by3(q[2]) << 2);
  if (key_mode == TRACE_KEY_M30) code >>= 33;
  keys[i] = code;
}

// ---- unique keys --------------------------------------------------------
constexpr int kUThreads = 256;
constexpr int kUPerThread = 16;
constexpr int kUTile = kUThreads * kUPerThread;

__device__ inline uint32_t block_excl_scan256(uint32_t v, uint32_t* s_wave, uint32_t* total) {
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
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i < w) off += s_wave[i];
    tot += s_wave[i];
  }
  __syncthreads();
  if (total) *total = tot;
  return off + x - v;
}

// Sorted position i starts a new leaf when its key differs from its
// predecessor's or (segmented meshes) its triangle belongs to another segment.
__device__ inline bool is_head(const uint64_t* __restrict__ keys, int64_t i,
                               const int32_t* __restrict__ tseg, const uint32_t* __restrict__ perm) {
  if (i == 0 || keys[i] != keys[i - 1]) return true;
  return tseg && tseg[perm[i]] != tseg[perm[i - 1]];
}

// A tile of kUTile sorted positions is read as kUPerThread rows of kUThreads
// consecutive positions (row j, thread t: position j * kUThreads + t), so every
// load of a row is one coalesced sweep.
__global__ __launch_bounds__(kUThreads) void count_heads_kernel(const uint64_t* __restrict__ keys,
                                                                int64_t n,
                                                                const int32_t* __restrict__ tseg,
                                                                const uint32_t* __restrict__ perm,
                                                                uint32_t* blk_cnt) {
  __shared__ uint32_t s_wave[4];
  const int64_t base = blockIdx.x * int64_t(kUTile) + threadIdx.x;
  uint32_t c = 0;
#pragma unroll 4
  for (int j = 0; j < kUPerThread; ++j) {
    const int64_t i = base + int64_t(j) * kUThreads;
    if (i < n) c += is_head(keys, i, tseg, perm) ? 1u : 0u;
  }
  uint32_t tot;
  block_excl_scan256(c, s_wave, &tot);
  if (threadIdx.x == 0) blk_cnt[blockIdx.x] = tot;
}

// Single workgroup: exclusive scan of the per-tile head counts; total -> header.
__global__ __launch_bounds__(256) void scan_blocks_kernel(uint32_t* blk, int64_t nblk,
                                                          BvhHeader* hdr) {
  __shared__ uint32_t s_wave[4];
  uint32_t carry = 0;
  for (int64_t b0 = 0; b0 < nblk; b0 += 256) {
    const int64_t b = b0 + threadIdx.x;
    const uint32_t v = b < nblk ? blk[b] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan256(v, s_wave, &tot);
    if (b < nblk) blk[b] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    hdr->n_leaves = int32_t(carry);
    hdr->n_inner = int32_t(carry) - 1;
    hdr->depth = 0;
  }
}

// Same row layout as count_heads_kernel.  Heads keep sorted order: the rank of
// the head at (row j, wave w, lane l) is the tile's offset, plus the heads of
// every earlier (row, wave) pair, plus the heads of lower lanes in its wave.
__global__ __launch_bounds__(kUThreads) void write_unique_kernel(
    const uint64_t* __restrict__ keys, int64_t n, const uint32_t* __restrict__ blk_off,
    const int32_t* __restrict__ tseg, const uint32_t* __restrict__ perm,
    uint64_t* __restrict__ ukeys, uint32_t* __restrict__ run_start, int32_t* __restrict__ useg,
    BvhHeader* hdr) {
  constexpr int kWaves = kUThreads / 64;
  constexpr int kPairs = kUPerThread * kWaves;
  static_assert(kPairs == 64, "one wave scans the (row, wave) head counts");
  __shared__ uint32_t s_off[kPairs];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t base = blockIdx.x * int64_t(kUTile) + threadIdx.x;
  uint32_t heads = 0;
#pragma unroll 4
  for (int j = 0; j < kUPerThread; ++j) {
    const int64_t i = base + int64_t(j) * kUThreads;
    const bool hd = i < n && is_head(keys, i, tseg, perm);
    if (hd) heads |= 1u << j;
    const uint64_t m = __ballot(hd);
    if (lane == 0) s_off[j * kWaves + w] = uint32_t(__popcll(m));
  }
  __syncthreads();
  if (w == 0) {
    // exclusive scan over the 64 (row, wave) counts, in row-major order
    const uint32_t v = s_off[lane];
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    s_off[lane] = x - v;
  }
  __syncthreads();
  const uint32_t tile0 = blk_off[blockIdx.x];
  const uint64_t below = (uint64_t(1) << lane) - 1u;
#pragma unroll 4
  for (int j = 0; j < kUPerThread; ++j) {
    const bool hd = (heads >> j) & 1u;
    const uint64_t m = __ballot(hd);
    if (hd) {
      const int64_t i = base + int64_t(j) * kUThreads;
      const uint32_t r = tile0 + s_off[j * kWaves + w] + uint32_t(__popcll(m & below));
      ukeys[r] = keys[i];
      run_start[r] = uint32_t(i);
      if (useg) useg[r] = tseg[perm[i]];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) run_start[hdr->n_leaves] = uint32_t(n);
}

// ---- hierarchy ------------------------------------------------------------
__device__ inline int delta_of(const uint64_t* k, int32_t u, int32_t i, int32_t j) {
  if (j < 0 || j >= u) return -1;
  return __clzll(k[i] ^ k[j]);
}

// Inner node i of the radix tree over u distinct keys (Karras 2012, Alg. 1).
// The lower key range becomes children[0], as lbvh.go:163-176 orders them.
// Forests (useg != null): unique key r belongs to segment s = useg[r], whose
// keys are [b, b + u) with b = u_begin[s]; the segment's tree is built over
// its own keys only (delta is -1 outside them) and its inner nodes take the
// indices b - s + [0, u - 1).
__global__ __launch_bounds__(256) void karras_kernel(const uint64_t* __restrict__ k_all,
                                                     const uint32_t* __restrict__ run_start_all,
                                                     const BvhHeader* __restrict__ hdr,
                                                     const int32_t* __restrict__ useg,
                                                     const int32_t* __restrict__ u_begin,
                                                     BvhNode* __restrict__ nodes_all,
                                                     int32_t* __restrict__ parent_inner_all,
                                                     int32_t* __restrict__ parent_leaf_all,
                                                     int32_t* __restrict__ inner_range_all,
                                                     uint32_t* __restrict__ counters) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= hdr->n_leaves) return;
  int32_t b = 0, s_id = 0, u = hdr->n_leaves;
  if (useg) {
    s_id = useg[r];
    b = u_begin[s_id];
    u = u_begin[s_id + 1] - b;
  }
  const int32_t i = r - b;
  if (i >= u - 1) {
    if (u == 1) parent_leaf_all[b] = -1;  // a one-leaf tree: its leaf is the root
    return;
  }
  const int32_t ib = b - s_id;  // first inner-node index of this tree
  counters[ib + i] = 0u;        // the refit's arrival counter of this inner node
  const uint64_t* k = k_all + b;
  const uint32_t* run_start = run_start_all + b;
  BvhNode* nodes = nodes_all + ib;
  int32_t* parent_inner = parent_inner_all + ib;
  int32_t* parent_leaf = parent_leaf_all + b;
  int32_t* inner_range = inner_range_all + 2 * ib;
  const int d = (delta_of(k, u, i, i + 1) - delta_of(k, u, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta_of(k, u, i, i - d);
  int32_t lmax = 2;
  while (delta_of(k, u, i, i + lmax * d) > dmin) lmax <<= 1;
  int32_t l = 0;
  for (int32_t t = lmax >> 1; t >= 1; t >>= 1)
    if (delta_of(k, u, i, i + (l + t) * d) > dmin) l += t;
  const int32_t j = i + l * d;
  const int dnode = delta_of(k, u, i, j);
  int32_t s = 0;
  int32_t t = l;
  do {
    t = (t + 1) >> 1;
    if (delta_of(k, u, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int32_t gamma = i + s * d + (d < 0 ? -1 : 0);
  const int32_t first = i < j ? i : j;
  const int32_t last = i < j ? j : i;
  inner_range[2 * i] = b + first;  // unique-key indices of the whole array
  inner_range[2 * i + 1] = b + last;

  const int32_t me = ib + i;  // global inner index
  BvhNode& nd = nodes[i];
  // child 0 covers [first, gamma], child 1 covers [gamma + 1, last]
  if (first == gamma) {
    nd.ref[0] = int32_t(run_start[gamma]);
    nd.cnt[0] = int32_t(run_start[gamma + 1] - run_start[gamma]);
    parent_leaf[gamma] = (me << 1) | 0;
  } else {
    nd.ref[0] = ib + gamma;
    nd.cnt[0] = 0;
    parent_inner[gamma] = (me << 1) | 0;
  }
  if (last == gamma + 1) {
    nd.ref[1] = int32_t(run_start[gamma + 1]);
    nd.cnt[1] = int32_t(run_start[gamma + 2] - run_start[gamma + 1]);
    parent_leaf[gamma + 1] = (me << 1) | 1;
  } else {
    nd.ref[1] = ib + gamma + 1;
    nd.cnt[1] = 0;
    parent_inner[gamma + 1] = (me << 1) | 1;
  }
  if (i == 0) parent_inner[0] = -1;  // the tree's root
}

// Forests: first unique key of each segment (u_begin[nseg] = total).
__global__ void seg_begin_kernel(const int32_t* __restrict__ useg, const BvhHeader* __restrict__ hdr,
                                 int32_t nseg, int32_t* __restrict__ u_begin) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t U = hdr->n_leaves;
  if (r == 0) u_begin[nseg] = U;
  if (r >= U) return;
  if (r == 0 || useg[r] != useg[r - 1]) u_begin[useg[r]] = r;
}

// FP32 triangles (v0, v1, v2) and outward-rounded FP32 boxes, in sorted
// order.  Vertices rather than edges are kept so that the winning hit of a
// ray can be re-evaluated in FP64 from the same vertices the reference uses.
// Gathers each sorted position's triangle; the FP32 triangles and boxes of a
// tile of kPosTile are staged in LDS and written with 16-byte stores.
__device__ inline void store_tile(const float* s, float* __restrict__ dst, int words) {
  float4* d4 = reinterpret_cast<float4*>(dst);  // tiles start on 16-byte boundaries
  const float4* s4 = reinterpret_cast<const float4*>(s);
  for (int k = threadIdx.x; k < words / 4; k += blockDim.x) d4[k] = s4[k];
  for (int k = (words & ~3) + int(threadIdx.x); k < words; k += blockDim.x) dst[k] = s[k];
}

__global__ __launch_bounds__(kPosTile) void reorder_kernel(const double* __restrict__ pos, int64_t n,
                                                           const uint32_t* __restrict__ perm,
                                                           float* __restrict__ tris, float* __restrict__ boxes) {
  __shared__ __align__(16) float s_tri[9 * kPosTile];
  __shared__ __align__(16) float s_box[6 * kPosTile];
  const int64_t base = blockIdx.x * int64_t(kPosTile);
  const int64_t i = base + threadIdx.x;
  if (i < n) {
    const uint32_t src = perm[i] < uint64_t(n) ? perm[i] : 0u;  // perm is a bijection on [0, n)
    const double* p = pos + 9 * int64_t(src);
    double v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = p[k];
    float* t = s_tri + 9 * threadIdx.x;
    float* b = s_box + 6 * threadIdx.x;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      t[a] = float(v[a]);
      t[3 + a] = float(v[3 + a]);
      t[6 + a] = float(v[6 + a]);
      double lo, hi;
      prim_bounds(v, a, lo, hi);
      b[a] = f32_down(lo);
      b[3 + a] = f32_up(hi);
    }
  }
  __syncthreads();
  const int cnt = n - base < kPosTile ? int(n - base) : kPosTile;
  store_tile(s_tri, tris + 9 * base, 9 * cnt);
  store_tile(s_box, boxes + 6 * base, 6 * cnt);
}

__device__ inline void st_f32(float* p, float v) {
  __hip_atomic_store(reinterpret_cast<uint32_t*>(p), __float_as_uint(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline float ld_f32(const float* p) {
  return __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT));
}

// Bottom-up refit: the leaf of distinct key u unions its triangles' boxes and
// writes the result into its parent's child slot; the second child to arrive
// at a node (per-node arrival counter, as bvh.go:206/219) unions both slots
// and continues upwards.
//
// Workgroup g handles leaves [256g, 256g + 256).  A node whose whole leaf range
// lies inside that span (most nodes: all but the top few levels) can only be
// reached by this workgroup, so its counter, child boxes and heights live in
// LDS, indexed by the node's own key index (a Karras node's index is one end
// of its range).  Nodes whose range crosses the span are met by two
// workgroups: their box words cross through agent-scope atomics whose
// completion is awaited before the global counter is bumped.  A climb is a
// chain of dependent loads, so each node's range and parent link are loaded
// one level ahead, while the arrival below them is still in flight.
constexpr int kRefitThreads = 256;

__global__ __launch_bounds__(kRefitThreads) void refit_kernel(
    const float* __restrict__ boxes, const uint32_t* __restrict__ run_start,
    const int32_t* __restrict__ parent_inner, const int32_t* __restrict__ parent_leaf,
    const int32_t* __restrict__ inner_range, BvhNode* nodes, uint32_t* counters, int32_t* heights,
    BvhHeader* hdr, uint32_t* d_err, const int32_t* __restrict__ useg, BvhHeader* seg_hdr,
    int32_t nseg) {
  __shared__ uint32_t l_cnt[kRefitThreads];
  __shared__ float l_box[kRefitThreads][2][6];
  __shared__ int32_t l_h[kRefitThreads][2];
  l_cnt[threadIdx.x] = 0u;
  __syncthreads();
  const int32_t u = hdr->n_leaves;
  const int32_t span0 = blockIdx.x * kRefitThreads;  // first leaf of this workgroup
  const int32_t leaf = span0 + threadIdx.x;
  if (leaf >= u) return;
  if (!useg) {
    if (u == 1) {
      hdr->root_ref = 0;
      hdr->root_cnt = int32_t(run_start[1]);
      hdr->depth = 0;
      return;
    }
    if (leaf == 0) {
      hdr->root_ref = 0;
      hdr->root_cnt = 0;
    }
  }
  const int32_t n_inner = u - nseg;  // every tree of the forest has one leaf more than inner nodes
  int32_t link = parent_leaf[leaf];
  int32_t f = 0, l = 0, par = -1;
  if (link >= 0 && (link >> 1) < n_inner) {
    f = inner_range[2 * (link >> 1)];
    l = inner_range[2 * (link >> 1) + 1];
    par = parent_inner[link >> 1];
  }
  float lo[3], hi[3];
  const uint32_t s = run_start[leaf], e = run_start[leaf + 1];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = boxes[6 * size_t(s) + a];
    hi[a] = boxes[6 * size_t(s) + 3 + a];
  }
  for (uint32_t k = s + 1; k < e; ++k) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], boxes[6 * size_t(k) + a]);
      hi[a] = fmaxf(hi[a], boxes[6 * size_t(k) + 3 + a]);
    }
  }
  int32_t h = 0;
  // A radix tree over 64-bit keys is at most 65 levels deep; the bound and
  // the index check only guard against a malformed hierarchy.
  for (int step = 0; link >= 0; ++step) {
    const int32_t node = link >> 1;
    const int slot = link & 1;
    if (step > 128 || node >= n_inner) {
      atomicOr(d_err, TRACE_ERR_TREE);
      return;
    }
    int32_t nf = 0, nl = 0, npar = -1;
    if (par >= 0 && (par >> 1) < n_inner) {
      nf = inner_range[2 * (par >> 1)];
      nl = inner_range[2 * (par >> 1) + 1];
      npar = parent_inner[par >> 1];
    }
    float* nb = reinterpret_cast<float*>(&nodes[node]);
    float* my = nb + 6 * slot;
    if (f >= span0 && l < span0 + kRefitThreads) {
      // local node: key index = node index (+ segment number in a forest)
      const int32_t key = node + (useg ? useg[f] : 0);
      if (key < f || key > l) {  // not a Karras node: the tree is malformed
        atomicOr(d_err, TRACE_ERR_TREE);
        return;
      }
      const int li = key - span0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        my[a] = lo[a];
        my[3 + a] = hi[a];
        l_box[li][slot][a] = lo[a];
        l_box[li][slot][3 + a] = hi[a];
      }
      l_h[li][slot] = h;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      const uint32_t prev = atomicAdd(&l_cnt[li], 1u);
      if (prev == 0) return;  // sibling not done yet; it will continue
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        lo[a] = fminf(lo[a], l_box[li][slot ^ 1][a]);
        hi[a] = fmaxf(hi[a], l_box[li][slot ^ 1][3 + a]);
      }
      const int32_t ho = l_h[li][slot ^ 1];
      h = 1 + (h > ho ? h : ho);
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        st_f32(my + a, lo[a]);
        st_f32(my + 3 + a, hi[a]);
      }
      __hip_atomic_store(heights + 2 * node + slot, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // The box and height words are agent-scope atomic stores (written
      // through to the device coherence point); waiting for their completion
      // before the arrival counter is bumped makes them visible to whichever
      // workgroup, on whichever XCD, arrives second.  A workgroup-scope fence
      // alone emits no wait for global memory, and the counter could overtake
      // the boxes.
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_s_waitcnt(0);
      const uint32_t prev = atomicAdd(counters + node, 1u);
      if (prev == 0) return;  // sibling not done yet; it will continue
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      const float* other = nb + 6 * (slot ^ 1);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        lo[a] = fminf(lo[a], ld_f32(other + a));
        hi[a] = fmaxf(hi[a], ld_f32(other + 3 + a));
      }
      const int32_t ho = __hip_atomic_load(heights + 2 * node + (slot ^ 1), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
      h = 1 + (h > ho ? h : ho);
    }
    link = par;
    f = nf;
    l = nl;
    par = npar;
  }
  // reached a root (inner node 0, or a forest tree's first inner node)
  if (useg) seg_hdr[useg[leaf]].depth = h;
  else hdr->depth = h;
}

// Forests: per-segment headers.  A segment with one distinct key is a root
// leaf.  n_inner and n_prims hold the forest totals: kernels use them only as
// bounds for node and triangle references.
__global__ void seg_finalize_kernel(const int32_t* __restrict__ u_begin,
                                    const uint32_t* __restrict__ run_start, int32_t nseg,
                                    BvhHeader* __restrict__ hdr, BvhHeader* __restrict__ seg_hdr) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) hdr->n_inner = hdr->n_leaves - nseg;  // the forest's inner nodes, all trees
  if (s >= nseg) return;
  const int32_t b = u_begin[s], e = u_begin[s + 1];
  BvhHeader& h = seg_hdr[s];
  h.n_leaves = e - b;
  h.n_inner = hdr->n_leaves - nseg;
  h.n_prims = hdr->n_prims;
  if (e - b == 1) {
    h.root_ref = int32_t(run_start[b]);
    h.root_cnt = int32_t(run_start[e] - run_start[b]);
    h.depth = 0;
  } else {
    h.root_ref = b - s;
    h.root_cnt = 0;
  }
}

static int grid_for(int64_t n, int threads) { return int(std::max<int64_t>(1, ceil_div(n, threads))); }

namespace {

// Scratch layout of build(): enclosing words, qbox, sort ping-pong buffers,
// sort status, unique keys, per-tile head counts, refit counters and heights.
struct LbvhPlan {
  size_t o_enc, o_qbox, o_kb, o_vb, o_st, o_uk, o_blk, o_cnt, o_hgt, o_useg, o_ub, total;
  LbvhPlan(int64_t n, int32_t nseg, int key_bits, int seg_passes, const onesweep::Config& cfg) {
    const bool forest = nseg > 1;
    const int64_t utiles = ceil_div(n, kUTile);
    ScratchPlan plan;
    o_enc = plan.take(6 * std::max<int64_t>(kEncBlocks, nseg) * sizeof(unsigned long long));
    o_qbox = plan.take(6 * size_t(nseg) * sizeof(double));
    o_kb = plan.take(n * sizeof(uint64_t));
    o_vb = plan.take(n * sizeof(uint32_t));
    o_st = plan.take(onesweep::status_words(n, key_bits, seg_passes, cfg) * sizeof(uint32_t));
    o_uk = plan.take(n * sizeof(uint64_t));
    o_blk = plan.take(utiles * sizeof(uint32_t));
    o_cnt = plan.take(std::max<int64_t>(1, n) * sizeof(uint32_t));
    o_hgt = plan.take(2 * std::max<int64_t>(1, n) * sizeof(int32_t));
    o_useg = plan.take(forest ? n * sizeof(int32_t) : 0);
    o_ub = plan.take(forest ? (size_t(nseg) + 1) * sizeof(int32_t) : 0);
    total = plan.off;
  }
};

int seg_passes_for(int32_t nseg) { return nseg <= 1 ? 0 : nseg <= 256 ? 1 : nseg <= 65536 ? 2 : 3; }

struct QBox {
  double v[6];
};

// Writes a caller's enclosing box passed by value (no host-memory copy).
__global__ void qbox_kernel(QBox b, double* __restrict__ qbox) {
  if (threadIdx.x < 6) qbox[threadIdx.x] = b.v[threadIdx.x];
}

}  // namespace

size_t scratch_bytes(const trace_ctx* ctx, const trace_mesh* mesh, int key_mode) {
  const onesweep::Config cfg{ctx->tune.sort_bits, ctx->tune.sort_kpt};
  const int key_bits = key_mode == TRACE_KEY_M30 ? 30 : 63;
  return LbvhPlan(mesh->ntri, mesh->nseg, key_bits, seg_passes_for(mesh->nseg), cfg).total;
}

int build(trace_ctx* ctx, const trace_mesh* mesh, const double* h_enclosing, int key_mode,
          trace_bvh* bvh, const Workspace* ws) {
  const int64_t n = mesh->ntri;
  hipStream_t st = ctx->stream;
  const int key_bits = key_mode == TRACE_KEY_M30 ? 30 : 63;
  const onesweep::Config sort_cfg{ctx->tune.sort_bits, ctx->tune.sort_kpt};
  const int32_t nseg = mesh->nseg;
  const bool forest = nseg > 1;
  const int seg_passes = seg_passes_for(nseg);
  if (forest && h_enclosing)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "LBVH: a segmented mesh takes no caller enclosing box");
  if (nseg > (1 << 24))
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "LBVH: %d segments exceed the 2^24 the sort's segment digits cover", nseg);
  const LbvhPlan plan(n, nseg, key_bits, seg_passes, sort_cfg);
  if (ws && ws->fixed() && plan.total > ws->bytes)  // refused before anything changes
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "LBVH: %lld triangles need %zu bytes of scratch, the fixed workspace holds %zu",
               static_cast<long long>(n), plan.total, ws->bytes);

  TRACE_ENSURE(ctx, ws, bvh->keys, n);
  TRACE_ENSURE(ctx, ws, bvh->perm, n);
  TRACE_ENSURE(ctx, ws, bvh->tris, 9 * n);
  TRACE_ENSURE(ctx, ws, bvh->prim_box, 6 * n);
  TRACE_ENSURE(ctx, ws, bvh->nodes, std::max<int64_t>(1, n - 1));
  TRACE_ENSURE(ctx, ws, bvh->parent, 2 * n);
  TRACE_ENSURE(ctx, ws, bvh->run_start, n + 1);
  TRACE_ENSURE(ctx, ws, bvh->inner_range, 2 * std::max<int64_t>(1, n - 1));
  TRACE_ENSURE(ctx, ws, bvh->header, 1);
  if (forest) TRACE_ENSURE(ctx, ws, bvh->seg_headers, nseg);
  bvh->n_prims = n;
  bvh->has_spheres = mesh->has_spheres;
  bvh->key_mode = key_mode;
  bvh->arity = 2;
  bvh->nseg = nseg;
  bvh->seg_off_h = mesh->seg_off_h;

  char* base = nullptr;
  int rc = workspace_base(ctx, ws, plan.total, &base, "LBVH");
  if (rc) return rc;
  auto* enc = reinterpret_cast<unsigned long long*>(base + plan.o_enc);
  auto* qbox = reinterpret_cast<double*>(base + plan.o_qbox);
  int32_t* useg = forest ? reinterpret_cast<int32_t*>(base + plan.o_useg) : nullptr;
  int32_t* u_begin = forest ? reinterpret_cast<int32_t*>(base + plan.o_ub) : nullptr;
  const int32_t* tseg = forest ? mesh->tri_seg.p : nullptr;
  auto* keys_b = reinterpret_cast<uint64_t*>(base + plan.o_kb);
  auto* vals_b = reinterpret_cast<uint32_t*>(base + plan.o_vb);
  auto* status = reinterpret_cast<uint32_t*>(base + plan.o_st);
  auto* ukeys = reinterpret_cast<uint64_t*>(base + plan.o_uk);
  auto* blk = reinterpret_cast<uint32_t*>(base + plan.o_blk);
  auto* counters = reinterpret_cast<uint32_t*>(base + plan.o_cnt);
  auto* heights = reinterpret_cast<int32_t*>(base + plan.o_hgt);
  const int64_t utiles = ceil_div(n, kUTile);

  if (h_enclosing && ws && ws->fixed()) {
    QBox b;
    for (int k = 0; k < 6; ++k) b.v[k] = h_enclosing[k];
    qbox_kernel<<<1, 64, 0, st>>>(b, qbox);
  } else if (h_enclosing) {
    HIP_TRY(ctx, hipMemcpyAsync(qbox, h_enclosing, 6 * sizeof(double), hipMemcpyHostToDevice, st));
  }
  if (!forest) {
    const int enc_blocks = int(std::min<int64_t>(grid_for(n, 256), kEncBlocks));
    enclosing_kernel<<<enc_blocks, 256, 0, st>>>(mesh->pos.p, n, enc);
    header_kernel<<<1, 256, 0, st>>>(enc, enc_blocks, bvh->header.p, qbox, h_enclosing ? 1 : 0,
                                    int32_t(n));
  } else {
    seg_enc_init_kernel<<<grid_for(6 * int64_t(nseg), 256), 256, 0, st>>>(enc, nseg);
    seg_enclosing_kernel<<<int(ceil_div(n, kSegTile)), 256, 0, st>>>(mesh->pos.p, n, tseg, enc);
    seg_header_kernel<<<grid_for(nseg, 256), 256, 0, st>>>(enc, nseg, bvh->seg_headers.p, qbox,
                                                          bvh->header.p, int32_t(n));
  }
  morton_kernel<<<grid_for(n, kPosTile), kPosTile, 0, st>>>(mesh->pos.p, n, qbox, tseg, key_mode, bvh->keys.p);

  // the values ping-pong between the permutation itself and scratch, so an
  // even number of passes leaves the sorted order in place
  auto* perm_u = reinterpret_cast<uint32_t*>(bvh->perm.p);
  const int where = onesweep::sort_pairs(st, bvh->keys.p, perm_u, keys_b, vals_b, n, key_bits,
                                         status, ctx->d_err, true, tseg, seg_passes, sort_cfg);
  uint64_t* sorted_keys = where ? keys_b : bvh->keys.p;
  if (where) {
    HIP_TRY(ctx, hipMemcpyAsync(bvh->keys.p, keys_b, n * sizeof(uint64_t),
                                hipMemcpyDeviceToDevice, st));
    HIP_TRY(ctx, hipMemcpyAsync(bvh->perm.p, vals_b, n * sizeof(uint32_t),
                                hipMemcpyDeviceToDevice, st));
    sorted_keys = bvh->keys.p;
  }

  const uint32_t* sperm = reinterpret_cast<const uint32_t*>(bvh->perm.p);
  count_heads_kernel<<<int(utiles), kUThreads, 0, st>>>(sorted_keys, n, tseg, sperm, blk);
  scan_blocks_kernel<<<1, 256, 0, st>>>(blk, utiles, bvh->header.p);
  write_unique_kernel<<<int(utiles), kUThreads, 0, st>>>(sorted_keys, n, blk, tseg, sperm, ukeys,
                                                         bvh->run_start.p, useg, bvh->header.p);
  if (forest) seg_begin_kernel<<<grid_for(n, 256), 256, 0, st>>>(useg, bvh->header.p, nseg, u_begin);
  int32_t* parent_inner = bvh->parent.p;
  int32_t* parent_leaf = bvh->parent.p + n;
  // every link the refit follows is written here: children's, and -1 at roots
  karras_kernel<<<grid_for(n, 256), 256, 0, st>>>(ukeys, bvh->run_start.p, bvh->header.p, useg,
                                                  u_begin, bvh->nodes.p, parent_inner,
                                                  parent_leaf, bvh->inner_range.p, counters);
  reorder_kernel<<<grid_for(n, kPosTile), kPosTile, 0, st>>>(mesh->pos.p, n, sperm, bvh->tris.p, bvh->prim_box.p);
  refit_kernel<<<grid_for(n, kRefitThreads), kRefitThreads, 0, st>>>(
      bvh->prim_box.p, bvh->run_start.p, parent_inner, parent_leaf, bvh->inner_range.p, bvh->nodes.p,
      counters, heights, bvh->header.p, ctx->d_err, useg, forest ? bvh->seg_headers.p : nullptr, nseg);
  if (forest)
    seg_finalize_kernel<<<grid_for(nseg, 256), 256, 0, st>>>(u_begin, bvh->run_start.p, nseg,
                                                            bvh->header.p, bvh->seg_headers.p);
  HIP_TRY(ctx, hipGetLastError());
  return wide::widen(ctx, bvh, parent_inner, ws, parent_leaf);
}

}  // namespace lbvh
