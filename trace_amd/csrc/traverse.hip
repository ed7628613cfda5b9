// Bulk closest-hit kernel: one ray per lane, rays streamed from HBM.
#include "traverse_dev.hpp"

namespace traverse {

constexpr int kThreads = 128;
constexpr int kStack = 64;  // LBVH depth <= 63 (one level per key bit)

template <bool COUNT>
__global__ __launch_bounds__(kThreads) void closest_hit_kernel(
    const BvhNode* __restrict__ nodes, const float* __restrict__ tris,
    const int32_t* __restrict__ perm, const BvhHeader* __restrict__ hdr_p,
    const float4* __restrict__ rays, int64_t n, int32_t* __restrict__ hit_out,
    float* __restrict__ t_out, int32_t* __restrict__ cnt_nodes, int32_t* __restrict__ cnt_tris,
    uint32_t* d_err) {
  const int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x;
  if (i >= n) return;
  const BvhHeader hdr = *hdr_p;
  const float4 a = rays[2 * i];
  const float4 b = rays[2 * i + 1];
  Ray ray{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  int32_t nn = 0, nt = 0;
  bool overflow = false, bad = false;
  Hit h = closest_hit_one<kStack, COUNT>(nodes, tris, hdr, ray, nn, nt, overflow, bad);
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
  const int64_t blocks = ceil_div(n, kThreads);
  if (d_nodes && d_tris) {
    closest_hit_kernel<true><<<int(blocks), kThreads, 0, ctx->stream>>>(
        bvh->nodes.p, bvh->tris.p, bvh->perm.p, bvh->header.p,
        reinterpret_cast<const float4*>(d_rays), n, d_hit, d_t, d_nodes, d_tris, ctx->d_err);
  } else {
    closest_hit_kernel<false><<<int(blocks), kThreads, 0, ctx->stream>>>(
        bvh->nodes.p, bvh->tris.p, bvh->perm.p, bvh->header.p,
        reinterpret_cast<const float4*>(d_rays), n, d_hit, d_t, nullptr, nullptr, ctx->d_err);
  }
  HIP_TRY(ctx, hipGetLastError());
  return TRACE_OK;
}

}  // namespace traverse
