// Bulk closest-hit kernel: one ray per lane, rays streamed from HBM.
#include "traverse_dev.hpp"

namespace traverse {

// THREADS: workgroup size (chosen by the autotuner, default 128).  STACK 64
// covers every LBVH (one level per key bit); deeper PHR trees use 256.
template <int THREADS, int STACK, bool COUNT>
__global__ __launch_bounds__(THREADS) void closest_hit_kernel(
    const BvhNode* __restrict__ nodes, const float* __restrict__ tris,
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
  Hit h = closest_hit_one<STACK, COUNT>(nodes, tris, hdr, ray, nn, nt, overflow, bad,
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

template <int THREADS, int STACK, bool COUNT>
static void launch(trace_ctx* ctx, const trace_bvh* bvh, const float4* rays, int64_t n,
                   int32_t id_offset, int32_t* d_hit, float* d_t, int32_t* d_nodes, int32_t* d_tris) {
  closest_hit_kernel<THREADS, STACK, COUNT><<<int(ceil_div(n, THREADS)), THREADS, 0, ctx->stream>>>(
      bvh->nodes_p(), bvh->tris_p(), bvh->perm_p(), bvh->header_p(), rays, n, id_offset, d_hit, d_t,
      d_nodes, d_tris, ctx->d_err);
}

template <int THREADS>
static void dispatch(trace_ctx* ctx, const trace_bvh* bvh, const float4* rays, int64_t n,
                     int32_t id_offset, int32_t* d_hit, float* d_t, int32_t* d_nodes,
                     int32_t* d_tris) {
  const bool deep = bvh->depth_bound() > 63;
  const bool count = d_nodes && d_tris;
  if (count) {
    if (deep) launch<THREADS, 256, true>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris);
    else launch<THREADS, 64, true>(ctx, bvh, rays, n, id_offset, d_hit, d_t, d_nodes, d_tris);
  } else {
    if (deep) launch<THREADS, 256, false>(ctx, bvh, rays, n, id_offset, d_hit, d_t, nullptr, nullptr);
    else launch<THREADS, 64, false>(ctx, bvh, rays, n, id_offset, d_hit, d_t, nullptr, nullptr);
  }
}

int closest_hit(trace_ctx* ctx, const trace_bvh* bvh, const float* d_rays, int64_t n,
                int32_t* d_hit, float* d_t, int32_t* d_nodes, int32_t* d_tris) {
  if (n == 0) return TRACE_OK;
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
