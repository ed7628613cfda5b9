// Bulk closest-hit kernel: one ray per lane, rays streamed from HBM.
#include "traverse_dev.hpp"

namespace traverse {

constexpr int kThreads = 128;

// STACK 64 covers every LBVH (one level per key bit); deeper PHR trees use 256.
template <int STACK, bool COUNT>
__global__ __launch_bounds__(kThreads) void closest_hit_kernel(
    const BvhNode* __restrict__ nodes, const float* __restrict__ tris,
    const int32_t* __restrict__ perm, const BvhHeader* __restrict__ hdr_p,
    const float4* __restrict__ rays, int64_t n, int32_t* __restrict__ hit_out,
    float* __restrict__ t_out, int32_t* __restrict__ cnt_nodes, int32_t* __restrict__ cnt_tris,
    uint32_t* d_err) {
  __shared__ int32_t s_stack[kLdsStack * kThreads];
  const int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x;
  if (i >= n) return;
  const BvhHeader hdr = *hdr_p;
  const float4 a = rays[2 * i];
  const float4 b = rays[2 * i + 1];
  Ray ray{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  int32_t nn = 0, nt = 0;
  bool overflow = false, bad = false;
  Hit h = closest_hit_one<STACK, COUNT>(nodes, tris, hdr, ray, nn, nt, overflow, bad,
                                        s_stack + threadIdx.x, kThreads);
  refine_hit_fp64(tris, ray, h);
  if (overflow) atomicOr(d_err, TRACE_ERR_STACK);
  if (bad) atomicOr(d_err, TRACE_ERR_TREE);
  hit_out[i] = h.pos >= 0 ? perm[h.pos] : -1;
  t_out[i] = h.t;
  if (COUNT) {
    cnt_nodes[i] = nn;
    cnt_tris[i] = nt;
  }
}

int closest_hit(trace_ctx* ctx, const trace_bvh* bvh, const float* d_rays, int64_t n,
                int32_t* d_hit, float* d_t, int32_t* d_nodes, int32_t* d_tris) {
  if (n == 0) return TRACE_OK;
  if (bvh->host_depth > 255)
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "closest hit: tree depth %d exceeds 255", bvh->host_depth);
  const int64_t blocks = ceil_div(n, kThreads);
  const auto* rays = reinterpret_cast<const float4*>(d_rays);
  hipStream_t st = ctx->stream;
  const bool deep = bvh->host_depth > 63;
  if (d_nodes && d_tris) {
    if (deep)
      closest_hit_kernel<256, true><<<int(blocks), kThreads, 0, st>>>(
          bvh->nodes.p, bvh->tris.p, bvh->perm.p, bvh->header.p, rays, n, d_hit, d_t, d_nodes,
          d_tris, ctx->d_err);
    else
      closest_hit_kernel<64, true><<<int(blocks), kThreads, 0, st>>>(
          bvh->nodes.p, bvh->tris.p, bvh->perm.p, bvh->header.p, rays, n, d_hit, d_t, d_nodes,
          d_tris, ctx->d_err);
  } else {
    if (deep)
      closest_hit_kernel<256, false><<<int(blocks), kThreads, 0, st>>>(
          bvh->nodes.p, bvh->tris.p, bvh->perm.p, bvh->header.p, rays, n, d_hit, d_t, nullptr,
          nullptr, ctx->d_err);
    else
      closest_hit_kernel<64, false><<<int(blocks), kThreads, 0, st>>>(
          bvh->nodes.p, bvh->tris.p, bvh->perm.p, bvh->header.p, rays, n, d_hit, d_t, nullptr,
          nullptr, ctx->d_err);
  }
  HIP_TRY(ctx, hipGetLastError());
  return TRACE_OK;
}

}  // namespace traverse
