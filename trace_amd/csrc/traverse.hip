// Bulk closest-hit kernel: one ray per lane, rays streamed from HBM.
#include <algorithm>

#include "instances.hpp"
#include "scan.hpp"
#include "traverse_dev.hpp"

namespace traverse {

// THREADS: workgroup size (chosen by the autotuner, default 128).  STACK 128
// covers the wide tree of any binary tree of depth <= 63 (at most 32 wide
// levels, 3 deferred children each, plus the split-leaf levels the widening
// pass sizes to fit); deeper PHR trees use 512.
template <int THREADS, int STACK, bool COUNT, bool SPH>
__global__ __launch_bounds__(THREADS) void closest_hit_kernel(
    const WideNode* __restrict__ wnodes, const float* __restrict__ tris,
    const int32_t* __restrict__ perm, const BvhHeader* __restrict__ hdr_p,
    const float4* __restrict__ rays, int64_t n, int32_t id_offset, int32_t* __restrict__ hit_out,
    float* __restrict__ t_out, int32_t* __restrict__ cnt_nodes, int32_t* __restrict__ cnt_tris,
    uint32_t* d_err) {
  __shared__ int32_t s_stack[kLdsStack * THREADS];
  const int64_t i = blockIdx.x * int64_t(THREADS) + threadIdx.x;
  if (i >= n) return;
  const BvhHeader hdr = *hdr_p;
  const float4 a = rays[2 * i];
  const float4 b = rays[2 * i + 1];
  Ray ray{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  int32_t nn = 0, nt = 0;
  bool overflow = false, bad = false;
  Hit h = closest_hit_wide<STACK, COUNT, SPH>(wnodes, tris, hdr, ray, nn, nt, overflow, bad,
                                         s_stack + threadIdx.x, THREADS);
  refine_hit_fp64(tris, ray, h);
  if (overflow) atomicOr(d_err, TRACE_ERR_STACK);
  if (bad) atomicOr(d_err, TRACE_ERR_TREE);
  hit_out[i] = h.pos >= 0 ? perm[h.pos] - id_offset : -1;
  t_out[i] = h.t;
  if (COUNT) {
    cnt_nodes[i] = nn;
    cnt_tris[i] = nt;
  }
}

// Persistent variant: a grid sized to the resident waves.  Each wave claims
// chunks of kChunk consecutive ray indices from a global counter (one atomic
// per chunk) and hands them to its lanes as their rays finish, so a wave keeps
// its lanes busy until the rays run out instead of waiting for its slowest
// ray.  A wave leaves the loop once the counter has passed n, its chunk is
// used up and its lanes are idle.  On 1M rays over the 1M-triangle PHR tree a
// chunk of 64 beat 32, 128, 256 and 512 (the tail of the last chunks is
// shorter; the counter is still not contended), and a refill at 16 idle lanes
// beat 8, 24 and 32.  With 8M rays traced grouped by origin the same pair wins:
// 4.94-4.95 ms against 5.00-5.04 (chunks of 32 or 128) and 4.97-5.10 (refill at
// 8 or 24).
constexpr uint32_t kChunk = 64;
constexpr int kRefillAt = 16;

// The FP64 re-solve of each winner is left to refine_kernel: without it the
// loop's register peak is 64 VGPRs, so eight waves fit on a SIMD (asked for
// where nothing spills: the plain, triangle-only variant); hit_out receives
// the winner's sorted position and t_out its FP32 distance.
template <int THREADS, int STACK, bool COUNT, bool SPH>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(COUNT || SPH ? 1 : 8, 8))) void closest_hit_persistent(
    const WideNode* __restrict__ wnodes, const float* __restrict__ tris,
    const int32_t* __restrict__ perm, const BvhHeader* __restrict__ hdr_p,
    const float4* __restrict__ rays, int64_t n, int32_t id_offset, int32_t* __restrict__ hit_out,
    float* __restrict__ t_out, int32_t* __restrict__ cnt_nodes, int32_t* __restrict__ cnt_tris,
    uint32_t* d_err, uint32_t* work, const uint32_t* __restrict__ ordered, const float4* __restrict__ orays,
    int32_t* __restrict__ ohit, float* __restrict__ ot) {
  __shared__ int32_t s_stack[kLdsStack * THREADS];
  if (ordered && *ordered) {  // trace the cell-major copy (ray_order) into its own result arrays
    rays = orays;
    hit_out = ohit;
    t_out = ot;
  }
  const BvhHeader hdr = *hdr_p;
  const uint32_t n_bound = uint32_t(hdr.n_wide);
  const int lane = __lane_id();
  const uint64_t lt = lanemask_lt();
  TravSpill<STACK> spill;
  TravStack<STACK> stack;
  stack.lds = lds_column(s_stack + threadIdx.x);
  stack.stride = THREADS;
  stack.spill = spill.v;
  WideState s;
  Ray ray{};
  int64_t rid = -1;
  bool active = false, overflow = false, bad = false;
  bool fetching = true;                  // the same in every lane of the wave
  int64_t pool_next = 0, pool_end = 0;   // this wave's unclaimed ray indices
  int32_t nn = 0, nt = 0;
  auto finish = [&]() {
    hit_out[rid] = s.h.pos;
    t_out[rid] = s.h.t;
    if (COUNT) {
      cnt_nodes[rid] = nn;
      cnt_tris[rid] = nt;
    }
  };
  bool pending = false;  // a finished ray whose result is not written yet
  while (true) {
    const uint64_t need = ballot64(!active);
    // Lanes whose ray is done wait until a quarter of the wave is idle: the
    // refinement, the result writes and the next ray's set-up then run for
    // many lanes at once instead of costing the whole wave for one lane.
    if (__popcll(need) >= kRefillAt || need == ~0ull) {
      if (pending) {
        finish();
        pending = false;
      }
      if (pool_next >= pool_end && fetching) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(work, kChunk);
        base = __shfl(base, 0, 64);
        if (int64_t(base) >= n) {
          fetching = false;
        } else {
          pool_next = base;
          pool_end = std::min<int64_t>(int64_t(base) + kChunk, n);
        }
      }
      const int64_t avail = pool_end - pool_next;
      if (avail > 0) {
        const int64_t rank = __popcll(need & lt);
        if (!active && rank < avail) {
          rid = pool_next + rank;
          const float4 ra = rays[2 * rid];
          const float4 rb = rays[2 * rid + 1];
          ray = Ray{ra.x, ra.y, ra.z, ra.w, rb.x, rb.y, rb.z, rb.w};
          nn = nt = 0;
          active = wide_begin<COUNT, SPH>(tris, hdr, ray, s, nt, bad);
          if (!active) finish();
        }
        pool_next += std::min<int64_t>(__popcll(need), avail);
      }
    }
    if (!ballot64(active)) {
      if (!fetching && pool_next >= pool_end) break;
      continue;
    }
    if (active) {
      active = wide_step<STACK, COUNT, SPH>(wnodes, tris, n_bound, hdr.n_prims, s, stack, nn, nt,
                                       overflow, bad);
      pending = !active;
    }
  }
  if (overflow) atomicOr(d_err, TRACE_ERR_STACK);
  if (bad) atomicOr(d_err, TRACE_ERR_TREE);
}

// After closest_hit_persistent: each winner re-solved in FP64
// (refine_hit_fp64) and its sorted position mapped to the caller's index.
// ordered (optional): set when the rays were traced as a cell-major copy,
// whose re-solve is refine_traced_kernel's.
__global__ __launch_bounds__(256) void refine_kernel(const float* __restrict__ tris, const int32_t* __restrict__ perm,
                                                     int32_t n_prims, const float4* __restrict__ rays, int64_t n,
                                                     int32_t id_offset, int32_t* __restrict__ hit_out,
                                                     float* __restrict__ t_out, uint32_t* d_err,
                                                     const uint32_t* __restrict__ ordered) {
  if (ordered && *ordered) return;
  for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
    const int32_t pos = hit_out[i];
    if (pos < 0) continue;  // a miss: -1 and the ray's tmax are final
    if (pos >= n_prims) {
      atomicOr(d_err, TRACE_ERR_TREE);
      hit_out[i] = -1;
      continue;
    }
    const float4 a = rays[2 * i];
    const float4 b = rays[2 * i + 1];
    const Ray ray{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    Hit h;
    h.pos = pos;
    h.t = t_out[i];
    h.t64 = h.t;
    h.u = h.v = 0.0f;
    h.front = true;
    refine_hit_fp64(tris, ray, h);
    hit_out[i] = perm[pos] - id_offset;
    t_out[i] = h.t;
  }
}

// ---- ray order ---------------------------------------------------------------
//
// Incoherent batches (origins spread over the scene, random directions) are
// traced grouped by origin: each origin falls in one of kCells cells (4 per
// axis of the tree's root box, numbered along a Morton curve), and the rays are
// copied into a cell-major batch (stable: a cell's rays keep the input order).
// At any moment the resident waves then trace rays of one or two cells, whose
// part of the tree stays in the L2 and Infinity caches, instead of rays from
// everywhere.  The persistent kernel traces the copy and the FP64 re-solve
// follows in the same order (coalesced loads and stores); a last pass gathers
// each caller ray's result from where it was traced.
//
// Measured on 8M rays over the 1M-triangle PHR tree with the rays reordered on
// the host: origin codes of 3, 6, 9 and 30 bits trace in 4.75, 4.63-4.74, 4.62
// and 4.53 ms against 5.30 ms in input order, so a few dozen cells give almost
// all of it.  On the device, rays gathered through a permutation inside the
// traversal (5.09 ms), and a copy written lane by lane from a 512-cell ranking
// with results scattered back (5.50 ms), lost most of that to uncoalesced
// accesses; here both directions move whole runs.
constexpr int kCellBits = 2;
constexpr int kCells = 1 << (3 * kCellBits);
constexpr int kOrderThreads = 256;
constexpr int kOrderWaves = kOrderThreads / 64;
constexpr int kOrderPerThread = 4;
constexpr int kOrderTile = kOrderThreads * kOrderPerThread;  // rays per workgroup
constexpr int64_t kOrderMin = int64_t(1) << 18;             // smaller batches are traced in input order

// box: the root box's low corner, then cells per unit length on each axis
// (load_root_box).
__device__ inline uint32_t origin_cell(const float* box, float4 o) {
  const float oc[3] = {o.x, o.y, o.z};
  uint32_t key = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    // fmaxf/fminf return the non-NaN operand: a NaN coordinate lands in cell 0
    const uint32_t q = uint32_t(fminf(fmaxf((oc[a] - box[a]) * box[3 + a], 0.0f), float((1 << kCellBits) - 1)));
#pragma unroll
    for (int b = 0; b < kCellBits; ++b) key |= ((q >> b) & 1u) << (3 * b + a);
  }
  return key;
}

__device__ inline void load_root_box(const BvhHeader* __restrict__ hdr, float* box) {
  if (threadIdx.x < 3) {
    const float lo = hdr->root_lo[threadIdx.x], ext = hdr->root_hi[threadIdx.x] - lo;
    box[threadIdx.x] = lo;
    box[3 + threadIdx.x] = ext > 0.0f ? float(1 << kCellBits) / ext : 0.0f;
  }
}

// The copy pays only when the origins are spread out: *ordered = 0 (trace in
// input order) when one cell holds more than half of an evenly strided sample
// of kProbe rays, as a camera's rays, which share one origin, all do.  One
// workgroup.
constexpr int kProbe = 1024;

__global__ __launch_bounds__(kOrderThreads) void cell_probe_kernel(const BvhHeader* __restrict__ hdr,
                                                                   const float4* __restrict__ rays, int64_t n,
                                                                   uint32_t* __restrict__ ordered) {
  __shared__ uint32_t h[kCells];
  __shared__ float box[6];
  for (int c = threadIdx.x; c < kCells; c += kOrderThreads) h[c] = 0u;
  load_root_box(hdr, box);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kProbe / kOrderThreads; ++k) {
    const int64_t i = int64_t(k * kOrderThreads + threadIdx.x) * n / kProbe;
    atomicAdd(&h[origin_cell(box, rays[2 * i])], 1u);
  }
  __syncthreads();
  bool crowded = false;
  for (int c = threadIdx.x; c < kCells; c += kOrderThreads) crowded |= 2 * h[c] > uint32_t(kProbe);
  crowded = __syncthreads_or(crowded);
  if (threadIdx.x == 0) *ordered = crowded ? 0u : 1u;
}

// counts[c * nblocks + b] = the rays of tile b in cell c (cell-major, so that
// one exclusive scan gives every tile's run start in every cell).
__global__ __launch_bounds__(kOrderThreads) void cell_count_kernel(const BvhHeader* __restrict__ hdr,
                                                                   const float4* __restrict__ rays, int64_t n,
                                                                   int32_t nblocks, uint32_t* __restrict__ counts,
                                                                   const uint32_t* __restrict__ ordered) {
  if (!*ordered) return;
  __shared__ uint32_t h[kCells];
  __shared__ float box[6];
  for (int c = threadIdx.x; c < kCells; c += kOrderThreads) h[c] = 0u;
  load_root_box(hdr, box);
  __syncthreads();
  const int64_t base = int64_t(blockIdx.x) * kOrderTile;
#pragma unroll
  for (int k = 0; k < kOrderPerThread; ++k) {
    const int64_t i = base + k * kOrderThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[origin_cell(box, rays[2 * i])], 1u);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < kCells; c += kOrderThreads) counts[int64_t(c) * nblocks + blockIdx.x] = h[c];
}

// The cell-major copy of one tile: its rays are ranked stably by cell (one
// ballot per cell bit in each wave, per-wave counts scanned in LDS), staged in
// LDS in cell order, and written as one run per cell at the cell's offset for
// this tile; inv[i] = where ray i went.
__global__ __launch_bounds__(kOrderThreads) void cell_scatter_kernel(const BvhHeader* __restrict__ hdr,
                                                                     const float4* __restrict__ rays, int64_t n,
                                                                     int32_t nblocks,
                                                                     const uint32_t* __restrict__ offsets,
                                                                     float4* __restrict__ sorted,
                                                                     uint32_t* __restrict__ inv,
                                                                     const uint32_t* __restrict__ ordered) {
  if (!*ordered) return;
  __shared__ float4 sray[2 * kOrderTile];
  __shared__ uint8_t scell[kOrderTile];
  __shared__ uint32_t slot[kOrderPerThread * kOrderWaves][kCells];  // per (k, wave): count, then prefix
  __shared__ uint32_t cstart[kCells];                               // cell's first local position
  __shared__ uint32_t coff[kCells];                                 // cell's run start in the copy
  __shared__ float box[6];
  const int tid = threadIdx.x, wave = tid / 64;
  for (int e = tid; e < kOrderPerThread * kOrderWaves * kCells; e += kOrderThreads) (&slot[0][0])[e] = 0u;
  load_root_box(hdr, box);
  __syncthreads();
  const int64_t base = int64_t(blockIdx.x) * kOrderTile;
  const uint64_t lt = lanemask_lt();
  uint32_t c[kOrderPerThread], r[kOrderPerThread];
#pragma unroll
  for (int k = 0; k < kOrderPerThread; ++k) {
    const int64_t i = base + k * kOrderThreads + tid;
    const bool valid = i < n;
    c[k] = valid ? origin_cell(box, rays[2 * i]) : 0u;
    // lanes of this wave with the same cell
    uint64_t m = ballot64(valid);
#pragma unroll
    for (int bit = 0; bit < 3 * kCellBits; ++bit) {
      const bool set = (c[k] >> bit) & 1u;
      const uint64_t bb = ballot64(set);
      m &= set ? bb : ~bb;
    }
    r[k] = uint32_t(__popcll(m & lt));
    if (valid && (m & lt) == 0u) slot[k * kOrderWaves + wave][c[k]] = uint32_t(__popcll(m));
  }
  __syncthreads();
  // per cell: exclusive prefix over the (k, wave) slots in ray order, then the
  // cells' starts inside the tile
  if (tid < kCells) {
    uint32_t run = 0;
    for (int sl = 0; sl < kOrderPerThread * kOrderWaves; ++sl) {
      const uint32_t v = slot[sl][tid];
      slot[sl][tid] = run;
      run += v;
    }
    cstart[tid] = run;  // the cell's total, scanned below
    coff[tid] = offsets[int64_t(tid) * nblocks + blockIdx.x];
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t acc = 0;
    for (int cc = 0; cc < kCells; ++cc) {
      const uint32_t v = cstart[cc];
      cstart[cc] = acc;
      acc += v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kOrderPerThread; ++k) {
    const int64_t i = base + k * kOrderThreads + tid;
    if (i < n) {
      const uint32_t within = slot[k * kOrderWaves + wave][c[k]] + r[k];
      const uint32_t p = cstart[c[k]] + within;
      // the ray is read again (its origin from L2): holding both halves of
      // every ray across the barriers would take 32 more VGPRs
      sray[2 * p] = rays[2 * i];
      sray[2 * p + 1] = rays[2 * i + 1];
      scell[p] = uint8_t(c[k]);
      inv[i] = coff[c[k]] + within;
    }
  }
  __syncthreads();
  const int valid_n = int(std::min<int64_t>(kOrderTile, n - base));
  for (int p = tid; p < valid_n; p += kOrderThreads) {
    const uint32_t cc = scell[p];
    const int64_t d = int64_t(coff[cc]) + (uint32_t(p) - cstart[cc]);
    sorted[2 * d] = sray[2 * p];
    sorted[2 * d + 1] = sray[2 * p + 1];
  }
}

// The re-solve of a cell-major batch, in traced order (coalesced, and the
// winners of neighbouring rays lie close together in the triangle array):
// res[j] = (caller's triangle index or -1, FP32 bits of t) of traced ray j.
__global__ __launch_bounds__(256) void refine_traced_kernel(const float* __restrict__ tris,
                                                            const int32_t* __restrict__ perm, int32_t n_prims,
                                                            const float4* __restrict__ rays, int64_t n,
                                                            int32_t id_offset, const int32_t* __restrict__ hit,
                                                            const float* __restrict__ t, int2* __restrict__ res,
                                                            uint32_t* d_err, const uint32_t* __restrict__ ordered) {
  if (!*ordered) return;
  for (int64_t j = blockIdx.x * int64_t(256) + threadIdx.x; j < n; j += int64_t(gridDim.x) * 256) {
    const int32_t pos = hit[j];
    if (pos < 0 || pos >= n_prims) {  // a miss: -1 and the ray's tmax are final
      if (pos >= n_prims) atomicOr(d_err, TRACE_ERR_TREE);
      res[j] = make_int2(-1, __float_as_int(t[j]));
      continue;
    }
    const float4 a = rays[2 * j];
    const float4 b = rays[2 * j + 1];
    const Ray ray{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    Hit h;
    h.pos = pos;
    h.t = t[j];
    h.t64 = h.t;
    h.u = h.v = 0.0f;
    h.front = true;
    refine_hit_fp64(tris, ray, h);
    res[j] = make_int2(perm[pos] - id_offset, __float_as_int(h.t));
  }
}

// Caller ray i's result from where it was traced (one 8-byte gather).
__global__ __launch_bounds__(256) void unpermute_kernel(const uint32_t* __restrict__ inv, const int2* __restrict__ res,
                                                        int64_t n, int32_t* __restrict__ hit_out,
                                                        float* __restrict__ t_out, const uint32_t* __restrict__ ordered) {
  if (!*ordered) return;
  for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
    const int2 r = res[inv[i]];
    hit_out[i] = r.x;
    t_out[i] = __int_as_float(r.y);
  }
}

// A cell-major batch in scratch: the rays, where each caller ray went, and room
// for the traced order's hits, distances and re-solved results.
struct RayOrder {
  const uint32_t* ordered = nullptr;  // device word: 1 = the copy below is traced
  const float4* rays = nullptr;
  const uint32_t* inv = nullptr;
  int32_t* hit = nullptr;
  float* t = nullptr;
  int2* res = nullptr;
};

static void order_plan(int64_t n, ScratchPlan& plan, size_t* o) {
  const int64_t nblocks = ceil_div(n, int64_t(kOrderTile));
  const int64_t ncount = nblocks * kCells;
  o[0] = plan.take(size_t(n) * 2 * sizeof(float4));           // sorted rays
  o[1] = plan.take(size_t(n) * sizeof(uint32_t));             // inv
  o[2] = plan.take(size_t(n) * sizeof(int32_t));              // traced hits
  o[3] = plan.take(size_t(n) * sizeof(float));                // traced distances
  o[4] = plan.take(size_t(ncount) * sizeof(uint32_t));        // counts
  o[5] = plan.take(size_t(ncount) * sizeof(uint32_t));        // offsets
  o[6] = plan.take(sizeof(uint32_t));                         // scan total
  o[7] = plan.take(scan::scratch_words(ncount) * sizeof(uint32_t));
  o[8] = plan.take(size_t(n) * sizeof(int2));                 // re-solved results
  o[9] = plan.take(sizeof(uint32_t));                         // ordered flag
}

// Enqueues the cell-major copy of n rays into scratch (made or skipped on the
// device, see cell_probe_kernel); out stays empty (trace in input order) for
// batches below kOrderMin, for calls that would have to grow the scratch arena
// while graphs hold it, and when the arena cannot grow.
static int ray_order(trace_ctx* ctx, const trace_bvh* bvh, const float4* rays, int64_t n, RayOrder* out) {
  *out = RayOrder{};
  if (!ctx->tune.trav_order || n < kOrderMin) return TRACE_OK;
  ScratchPlan plan;
  size_t o[10];
  order_plan(n, plan, o);
  if (plan.off > ctx->scratch_bytes && (ctx->capturing || ctx->live_graphs > 0)) return TRACE_OK;
  if (ctx_reserve_scratch(ctx, plan.off) != TRACE_OK) {
    // no room for the copy (46 bytes a ray): trace in input order; the failed
    // allocation is not this call's error
    (void)hipGetLastError();
    return TRACE_OK;
  }
  char* base = static_cast<char*>(ctx->scratch);
  auto* sorted = reinterpret_cast<float4*>(base + o[0]);
  auto* inv = reinterpret_cast<uint32_t*>(base + o[1]);
  auto* counts = reinterpret_cast<uint32_t*>(base + o[4]);
  auto* offsets = reinterpret_cast<uint32_t*>(base + o[5]);
  auto* total = reinterpret_cast<uint32_t*>(base + o[6]);
  auto* ordered = reinterpret_cast<uint32_t*>(base + o[9]);
  const int32_t nblocks = int32_t(ceil_div(n, int64_t(kOrderTile)));
  cell_probe_kernel<<<1, kOrderThreads, 0, ctx->stream>>>(bvh->header_p(), rays, n, ordered);
  cell_count_kernel<<<nblocks, kOrderThreads, 0, ctx->stream>>>(bvh->header_p(), rays, n, nblocks, counts, ordered);
  // (with the copy skipped, this scans stale counts that nothing reads)
  scan::exclusive_u32(ctx->stream, counts, offsets, int64_t(nblocks) * kCells, total,
                      reinterpret_cast<uint32_t*>(base + o[7]));
  cell_scatter_kernel<<<nblocks, kOrderThreads, 0, ctx->stream>>>(bvh->header_p(), rays, n, nblocks, offsets, sorted,
                                                                  inv, ordered);
  *out = RayOrder{ordered, sorted, inv, reinterpret_cast<int32_t*>(base + o[2]), reinterpret_cast<float*>(base + o[3]),
                  reinterpret_cast<int2*>(base + o[8])};
  return TRACE_OK;
}

template <int THREADS, int STACK, bool COUNT, bool SPH>
static int launch_kind(trace_ctx* ctx, const trace_bvh* bvh, const float4* rays, int64_t n,
                       int32_t id_offset, int32_t* d_hit, float* d_t, int32_t* d_nodes, int32_t* d_tris) {
  if (ctx->tune.trav_persistent) {
    // counting mode keeps the input order (its per-ray counts are written by the traversal)
    RayOrder ro;
    if (!COUNT) {
      const int rc = ray_order(ctx, bvh, rays, n, &ro);
      if (rc) return rc;
    }
    // up to 32 resident waves per CU; workgroups beyond what fits start as
    // others leave and find the counter exhausted
    const int64_t blocks = std::min<int64_t>(ceil_div(n, THREADS), int64_t(ctx->n_cu) * 32 * 64 / THREADS);
    uint32_t* work = ctx->d_work;
    fill_async<uint32_t>(ctx->stream, work, 1, 0u);
    // with a copy made, ro.ordered decides on the device which arrays are traced
    closest_hit_persistent<THREADS, STACK, COUNT, SPH><<<int(blocks), THREADS, 0, ctx->stream>>>(
        bvh->wnodes_p(), bvh->tris_p(), bvh->perm_p(), bvh->header_p(), rays, n,
        id_offset, d_hit, d_t, d_nodes, d_tris, ctx->d_err, work, ro.ordered, ro.rays, ro.hit, ro.t);
    // sorted positions index the whole triangle array (a forest's, for a segment view)
    const int64_t n_sorted = bvh->base ? bvh->base->n_prims : bvh->n_prims;
    // grid-stride passes: the one of these that the device-side decision
    // leaves idle costs a few workgroups, not one per 256 rays
    const int grid = int(std::min<int64_t>(ceil_div(n, 256), int64_t(ctx->n_cu) * 32));
    refine_kernel<<<grid, 256, 0, ctx->stream>>>(bvh->tris_p(), bvh->perm_p(), int32_t(n_sorted), rays, n, id_offset,
                                                 d_hit, d_t, ctx->d_err, ro.ordered);
    if (ro.ordered) {
      refine_traced_kernel<<<grid, 256, 0, ctx->stream>>>(bvh->tris_p(), bvh->perm_p(), int32_t(n_sorted), ro.rays,
                                                          n, id_offset, ro.hit, ro.t, ro.res, ctx->d_err, ro.ordered);
      unpermute_kernel<<<grid, 256, 0, ctx->stream>>>(ro.inv, ro.res, n, d_hit, d_t, ro.ordered);
    }
    return TRACE_OK;
  }
  closest_hit_kernel<THREADS, STACK, COUNT, SPH><<<int(ceil_div(n, THREADS)), THREADS, 0, ctx->stream>>>(
      bvh->wnodes_p(), bvh->tris_p(), bvh->perm_p(), bvh->header_p(), rays, n, id_offset, d_hit, d_t,
      d_nodes, d_tris, ctx->d_err);
  return TRACE_OK;
}

template <int THREADS, int STACK, bool COUNT>
static int launch(trace_ctx* ctx, const trace_bvh* bvh, const float4* rays, int64_t n,
                  int32_t id_offset, int32_t* d_hit, float* d_t, int32_t* d_nodes, int32_t* d_tris) {
  if (bvh->spheres())
    return launch_kind<THREADS, STACK, COUNT, true>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris);
  return launch_kind<THREADS, STACK, COUNT, false>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris);
}

template <int THREADS>
static int dispatch(trace_ctx* ctx, const trace_bvh* bvh, const float4* rays, int64_t n,
                    int32_t id_offset, int32_t* d_hit, float* d_t, int32_t* d_nodes,
                    int32_t* d_tris) {
  const bool deep = bvh->depth_bound() > 63;
  const bool count = d_nodes && d_tris;
  if (count) {
    if (deep) return launch<THREADS, 512, true>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris);
    return launch<THREADS, 128, true>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris);
  }
  if (deep) return launch<THREADS, 512, false>(ctx, bvh, rays, n, id_offset, d_hit, d_t, nullptr, nullptr);
  return launch<THREADS, 128, false>(ctx, bvh, rays, n, id_offset, d_hit, d_t, nullptr, nullptr);
}

int closest_hit(trace_ctx* ctx, const trace_bvh* bvh, const float* d_rays, int64_t n,
                int32_t* d_hit, float* d_t, int32_t* d_nodes, int32_t* d_tris) {
  if (n == 0) return TRACE_OK;
  // persistent lanes claim rays through a 32-bit counter, and hit ids are int32
  if (n > int64_t(INT32_MAX))
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "closest hit: %lld rays exceed the 2^31 - 1 one call takes",
               static_cast<long long>(n));
  if (bvh->instanced) {
    if (d_nodes || d_tris) TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "closest hit: no counting mode on instanced trees");
    return instances::closest_hit(ctx, bvh, d_rays, n, d_hit, d_t);
  }
  if (bvh->depth_bound() > 255)
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "closest hit: tree depth %d exceeds 255", bvh->depth_bound());
  if (!bvh->base && bvh->nseg > 1)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "closest hit: a forest is queried through its segment views");
  const int32_t id_offset = bvh->base ? int32_t(bvh->base->seg_off_h[bvh->seg]) : 0;
  const auto* rays = reinterpret_cast<const float4*>(d_rays);
  int rc;
  switch (ctx->tune.trav_threads) {
    case 64: rc = dispatch<64>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris); break;
    case 256: rc = dispatch<256>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris); break;
    default: rc = dispatch<128>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris); break;
  }
  if (rc) return rc;
  HIP_TRY(ctx, hipGetLastError());
  return TRACE_OK;
}

}  // namespace traverse
