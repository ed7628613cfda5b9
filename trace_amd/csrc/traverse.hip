// Bulk closest-hit kernel: one ray per lane, rays streamed from HBM.
#include <algorithm>

#include "instances.hpp"
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
// beat 8, 24 and 32.
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
    uint32_t* d_err, uint32_t* work) {
  __shared__ int32_t s_stack[kLdsStack * THREADS];
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
__global__ __launch_bounds__(256) void refine_kernel(const float* __restrict__ tris, const int32_t* __restrict__ perm,
                                                     int32_t n_prims, const float4* __restrict__ rays, int64_t n,
                                                     int32_t id_offset, int32_t* __restrict__ hit_out,
                                                     float* __restrict__ t_out, uint32_t* d_err) {
  const int64_t i = blockIdx.x * int64_t(256) + threadIdx.x;
  if (i >= n) return;
  const int32_t pos = hit_out[i];
  if (pos < 0) return;  // a miss: -1 and the ray's tmax are final
  if (pos >= n_prims) {
    atomicOr(d_err, TRACE_ERR_TREE);
    hit_out[i] = -1;
    return;
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

template <int THREADS, int STACK, bool COUNT, bool SPH>
static void launch_kind(trace_ctx* ctx, const trace_bvh* bvh, const float4* rays, int64_t n,
                        int32_t id_offset, int32_t* d_hit, float* d_t, int32_t* d_nodes, int32_t* d_tris) {
  if (ctx->tune.trav_persistent) {
    // up to 32 resident waves per CU; workgroups beyond what fits start as
    // others leave and find the counter exhausted
    const int64_t blocks = std::min<int64_t>(ceil_div(n, THREADS), int64_t(ctx->n_cu) * 32 * 64 / THREADS);
    uint32_t* work = ctx->d_work;
    fill_async<uint32_t>(ctx->stream, work, 1, 0u);
    closest_hit_persistent<THREADS, STACK, COUNT, SPH><<<int(blocks), THREADS, 0, ctx->stream>>>(
        bvh->wnodes_p(), bvh->tris_p(), bvh->perm_p(), bvh->header_p(), rays, n,
        id_offset, d_hit, d_t, d_nodes, d_tris, ctx->d_err, work);
    // sorted positions index the whole triangle array (a forest's, for a segment view)
    const int64_t n_sorted = bvh->base ? bvh->base->n_prims : bvh->n_prims;
    refine_kernel<<<int(ceil_div(n, 256)), 256, 0, ctx->stream>>>(bvh->tris_p(), bvh->perm_p(), int32_t(n_sorted),
                                                                 rays, n, id_offset, d_hit, d_t, ctx->d_err);
    return;
  }
  closest_hit_kernel<THREADS, STACK, COUNT, SPH><<<int(ceil_div(n, THREADS)), THREADS, 0, ctx->stream>>>(
      bvh->wnodes_p(), bvh->tris_p(), bvh->perm_p(), bvh->header_p(), rays, n, id_offset, d_hit, d_t,
      d_nodes, d_tris, ctx->d_err);
}

// Trees holding Spheres take the variant whose leaf loop checks each
// primitive's tag; triangle-only trees keep the plain loop.
template <int THREADS, int STACK, bool COUNT>
static void launch(trace_ctx* ctx, const trace_bvh* bvh, const float4* rays, int64_t n,
                   int32_t id_offset, int32_t* d_hit, float* d_t, int32_t* d_nodes, int32_t* d_tris) {
  if (bvh->spheres())
    launch_kind<THREADS, STACK, COUNT, true>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris);
  else
    launch_kind<THREADS, STACK, COUNT, false>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris);
}

template <int THREADS>
static void dispatch(trace_ctx* ctx, const trace_bvh* bvh, const float4* rays, int64_t n,
                     int32_t id_offset, int32_t* d_hit, float* d_t, int32_t* d_nodes,
                     int32_t* d_tris) {
  const bool deep = bvh->depth_bound() > 63;
  const bool count = d_nodes && d_tris;
  if (count) {
    if (deep) launch<THREADS, 512, true>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris);
    else launch<THREADS, 128, true>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris);
  } else {
    if (deep) launch<THREADS, 512, false>(ctx, bvh, rays, n, id_offset, d_hit, d_t, nullptr, nullptr);
    else launch<THREADS, 128, false>(ctx, bvh, rays, n, id_offset, d_hit, d_t, nullptr, nullptr);
  }
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
  switch (ctx->tune.trav_threads) {
    case 64: dispatch<64>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris); break;
    case 256: dispatch<256>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris); break;
    default: dispatch<128>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris); break;
  }
  HIP_TRY(ctx, hipGetLastError());
  return TRACE_OK;
}

}  // namespace traverse
