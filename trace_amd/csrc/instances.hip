// Two-level trees for instanced geometry.  No reference counterpart: the
// reference's Scene.collectTracables copies the triangles of every placement
// of a mesh into one list (scene.go:100-114), which the large configuration's
// 5,000 placements of one 400-triangle sphere turn into 2M triangles.
//
// A bottom-level tree (BLAS) is the LBVH or PHR tree of one mesh in its own
// coordinates; an instance places a BLAS with an affine transform; the top
// level (TLAS) is an LBVH over the instances' world boxes, built by the
// ordinary LBVH build from one box-shaped primitive per instance.  A ray walks
// the TLAS in world space with the wide-node step of the closest-hit kernels;
// at an instance leaf it is moved into the instance's object space
// (o' = M o + c, d' = M d, with M, c the inverse transform, evaluated in FP64)
// and walks the BLAS bounded by the best distance so far.  Distances are the
// same in both spaces (the direction is transformed with the point), so the
// nearest hit over all instances wins, ties going to the larger flattened
// index.  The winner is re-solved in FP64 with the ray moved into its
// instance's space in FP64.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "instances.hpp"
#include "lbvh.hpp"
#include "traverse_dev.hpp"

namespace instances {
namespace {

struct InstRec {
  double w2o[12];    // world -> object, rows (M | c): object = M world + c
  int32_t blas;      // bottom-level tree
  int32_t tri_base;  // flattened index of the instance's first triangle
  int32_t pad[6];
};
static_assert(sizeof(InstRec) == 128, "InstRec is 128 bytes");

struct BlasDesc {
  const WideNode* wnodes;
  const float* tris;
  const int32_t* perm;
  const BvhHeader* hdr;
  int32_t id_offset;  // a forest view's first triangle (its perm holds forest indices)
  int32_t pad[7];
};
static_assert(sizeof(BlasDesc) == 64, "BlasDesc is 64 bytes");

__device__ inline traverse::Ray to_object(const InstRec& I, const traverse::Ray& w, float tmax) {
#pragma clang fp contract(off)
  const double* m = I.w2o;
  const double ox = w.ox, oy = w.oy, oz = w.oz, dx = w.dx, dy = w.dy, dz = w.dz;
  traverse::Ray r;
  r.ox = float(m[0] * ox + m[1] * oy + m[2] * oz + m[3]);
  r.oy = float(m[4] * ox + m[5] * oy + m[6] * oz + m[7]);
  r.oz = float(m[8] * ox + m[9] * oy + m[10] * oz + m[11]);
  r.dx = float(m[0] * dx + m[1] * dy + m[2] * dz);
  r.dy = float(m[4] * dx + m[5] * dy + m[6] * dz);
  r.dz = float(m[8] * dx + m[9] * dy + m[10] * dz);
  r.tmin = w.tmin;
  r.tmax = tmax;
  return r;
}

// The winner re-solved in FP64 (the operation order of refine_hit_fp64,
// primitive.go:172-208) with the ray moved into object space in FP64 and not
// rounded: the FP32 object-space ray of the search is off by an FP32 rounding
// of the moved origin, which a shrinking transform and a short hit magnify.
__device__ inline double refine_t_fp64(const float* __restrict__ tris, int32_t pos, const InstRec& I,
                                       const traverse::Ray& w, double t_fallback) {
#pragma clang fp contract(off)
  const double* m = I.w2o;
  const double wx = w.ox, wy = w.oy, wz = w.oz, wdx = w.dx, wdy = w.dy, wdz = w.dz;
  const double ox = m[0] * wx + m[1] * wy + m[2] * wz + m[3];
  const double oy = m[4] * wx + m[5] * wy + m[6] * wz + m[7];
  const double oz = m[8] * wx + m[9] * wy + m[10] * wz + m[11];
  const double dx = m[0] * wdx + m[1] * wdy + m[2] * wdz;
  const double dy = m[4] * wdx + m[5] * wdy + m[6] * wdz;
  const double dz = m[8] * wdx + m[9] * wdy + m[10] * wdz;
  const float* t = tris + 9 * size_t(pos);
  const double v0x = t[0], v0y = t[1], v0z = t[2];
  const double e1x = double(t[3]) - v0x, e1y = double(t[4]) - v0y, e1z = double(t[5]) - v0z;
  const double e2x = double(t[6]) - v0x, e2y = double(t[7]) - v0y, e2z = double(t[8]) - v0z;
  const double px = dy * e2z - dz * e2y;
  const double py = dz * e2x - dx * e2z;
  const double pz = dx * e2y - dy * e2x;
  const double det = e1x * px + e1y * py + e1z * pz;
  if (det == 0.0) return t_fallback;
  const double inv = 1.0 / det;
  const double tx = ox - v0x, ty = oy - v0y, tz = oz - v0z;
  const double qx = ty * e1z - tz * e1y;
  const double qy = tz * e1x - tx * e1z;
  const double qz = tx * e1y - ty * e1x;
  return (e2x * qx + e2y * qy + e2z * qz) * inv;
}

// One ray per lane: the TLAS walk (stack in LDS columns s_top) calls a BLAS
// walk (its own stack, s_bot) at every instance leaf it enters.
template <int THREADS, int STACK>
__global__ __launch_bounds__(THREADS) void instanced_hit_kernel(
    const WideNode* __restrict__ tlas, const int32_t* __restrict__ tlas_perm, const BvhHeader* __restrict__ hdr_p,
    const InstRec* __restrict__ inst, const BlasDesc* __restrict__ blas, const float4* __restrict__ rays, int64_t n,
    int32_t* __restrict__ hit_out, float* __restrict__ t_out, uint32_t* d_err) {
  __shared__ int32_t s_top[traverse::kLdsStack * THREADS];
  __shared__ int32_t s_bot[traverse::kLdsStack * THREADS];
  const int64_t i = blockIdx.x * int64_t(THREADS) + threadIdx.x;
  if (i >= n) return;
  const BvhHeader th = *hdr_p;
  const float4 a = rays[2 * i];
  const float4 b = rays[2 * i + 1];
  const traverse::Ray ray{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  traverse::WideState s;
  traverse::TravSpill<STACK> spill;
  traverse::TravStack<STACK> stack;
  stack.lds = traverse::lds_column(s_top + threadIdx.x);
  stack.stride = THREADS;
  stack.spill = spill.v;
  bool overflow = false, bad = false;
  int32_t nn = 0, nt = 0;
  // s.h.t: the best distance so far; s.h.pos: the winner's flattened index
  int32_t best_inst = -1, best_local = -1;
  // an instance leaf's primitive k: the walk of its bottom-level tree
  auto prim = [&](int32_t k) {
    const int32_t q = tlas_perm[k];
    const InstRec& I = inst[q];
    const BlasDesc B = blas[I.blas];
    const BvhHeader bh = *B.hdr;
    const traverse::Ray ro = to_object(I, ray, s.h.t);
    const traverse::Hit h = traverse::closest_hit_wide<STACK, false, false>(
        B.wnodes, B.tris, bh, ro, nn, nt, overflow, bad, s_bot + threadIdx.x, THREADS);
    if (h.pos < 0) return;
    const int32_t gid = I.tri_base + (B.perm[h.pos] - B.id_offset);
    if (h.t < s.h.t || (h.t == s.h.t && gid > s.h.pos)) {
      s.h.t = h.t;
      s.h.pos = gid;
      best_inst = q;
      best_local = h.pos;
    }
  };
  s.r = traverse::ray_pre(ray);
  s.h.t = ray.tmax;
  s.h.t64 = ray.tmax;
  s.h.pos = -1;
  s.h.u = s.h.v = 0.0f;
  s.h.front = true;
  s.sp = 0;
  s.visits = 0;
  s.node = th.wroot;
  float tn;
  if (traverse::slab(s.r, th.root_lo[0], th.root_lo[1], th.root_lo[2], th.root_hi[0], th.root_hi[1], th.root_hi[2],
                     s.h.t, tn)) {
    if (th.root_cnt > 0 && th.wroot < 0) {
      // a top level of one leaf
      if (th.root_ref < 0 || th.root_cnt > th.n_prims - th.root_ref) bad = true;
      else
        for (int32_t k = th.root_ref; k < th.root_ref + th.root_cnt; ++k) prim(k);
    } else {
      while (traverse::wide_step_with<STACK, false>(tlas, uint32_t(th.n_wide), th.n_prims, s, stack, nn, overflow,
                                                    bad, prim)) {
      }
    }
  }
  float t = s.h.t;
  if (best_inst >= 0) {  // the winner in FP64, in its instance's space
    const InstRec& I = inst[best_inst];
    const BlasDesc B = blas[I.blas];
    t = float(refine_t_fp64(B.tris, best_local, I, ray, double(s.h.t)));
  }
  hit_out[i] = best_inst >= 0 ? s.h.pos : -1;
  t_out[i] = best_inst >= 0 ? t : ray.tmax;
  if (overflow) atomicOr(d_err, TRACE_ERR_STACK);
  if (bad) atomicOr(d_err, TRACE_ERR_TREE);
}

}  // namespace

int closest_hit(trace_ctx* ctx, const trace_bvh* bvh, const float* d_rays, int64_t n, int32_t* d_hit, float* d_t) {
  if (n == 0) return TRACE_OK;
  constexpr int kThreads = 64;
  instanced_hit_kernel<kThreads, 128><<<int(ceil_div(n, kThreads)), kThreads, 0, ctx->stream>>>(
      bvh->wnodes.p, bvh->perm.p, bvh->header.p, reinterpret_cast<const InstRec*>(bvh->inst.p),
      reinterpret_cast<const BlasDesc*>(bvh->blas.p), reinterpret_cast<const float4*>(d_rays), n, d_hit, d_t,
      ctx->d_err);
  HIP_TRY(ctx, hipGetLastError());
  return TRACE_OK;
}

}  // namespace instances

extern "C" {

int trace_instances_build(trace_ctx* ctx, const trace_bvh* const* h_blas, int32_t n_blas, const int32_t* h_inst_blas,
                          const double* h_xform, const int64_t* h_tri_base, int32_t n_inst, trace_bvh** out) {
  using namespace instances;
  if (!ctx) return TRACE_E_INVALID;
  if (!h_blas || !h_inst_blas || !h_xform || !h_tri_base || !out || n_blas < 1 || n_inst < 1)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_instances_build: null argument or nothing to place");
  *out = nullptr;
  std::vector<BlasDesc> descs(n_blas);
  std::vector<BvhHeader> bhdr(n_blas);
  int32_t depth = 0;
  for (int32_t j = 0; j < n_blas; ++j) {
    const trace_bvh* b = h_blas[j];
    if (!b) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_instances_build: bottom-level tree %d is null", j);
    if (b->instanced || (!b->base && b->nseg > 1))
      TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_instances_build: tree %d is instanced or a whole forest", j);
    if (b->spheres())
      TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_instances_build: tree %d holds spheres (triangles only)", j);
    if (b->depth_bound() > 63)
      TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_instances_build: tree %d is deeper than 63 levels", j);
    depth = std::max(depth, b->depth_bound());
    std::memset(&descs[j], 0, sizeof(BlasDesc));
    descs[j].wnodes = b->wnodes_p();
    descs[j].tris = b->tris_p();
    descs[j].perm = b->perm_p();
    descs[j].hdr = b->header_p();
    descs[j].id_offset = b->base ? int32_t(b->base->seg_off_h[b->seg]) : 0;
    HIP_TRY(ctx, hipMemcpyAsync(&bhdr[j], b->header_p(), sizeof(BvhHeader), hipMemcpyDeviceToHost, ctx->stream));
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<InstRec> recs(n_inst);
  std::vector<double> pos(9 * size_t(n_inst));
  int64_t flat = 0;
  for (int32_t i = 0; i < n_inst; ++i) {
    const int32_t j = h_inst_blas[i];
    if (j < 0 || j >= n_blas) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_instances_build: instance %d names tree %d", i, j);
    const int64_t nb = h_blas[j]->base ? h_blas[j]->base->seg_off_h[h_blas[j]->seg + 1] - descs[j].id_offset
                                       : h_blas[j]->n_prims;
    if (h_tri_base[i] < 0 || h_tri_base[i] + nb > int64_t(INT32_MAX))
      TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_instances_build: instance %d numbered outside 0 .. 2^31 - 1", i);
    flat = std::max(flat, h_tri_base[i] + nb);
    const double* x = h_xform + 12 * size_t(i);  // world = A object + t, rows (A | t)
    const double A[9] = {x[0], x[1], x[2], x[4], x[5], x[6], x[8], x[9], x[10]};
    const double t[3] = {x[3], x[7], x[11]};
    const double c00 = A[4] * A[8] - A[5] * A[7], c01 = A[5] * A[6] - A[3] * A[8], c02 = A[3] * A[7] - A[4] * A[6];
    const double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
    bool finite = std::isfinite(det);
    for (int k = 0; k < 12; ++k) finite = finite && std::isfinite(x[k]);
    if (!finite || det == 0.0)
      TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_instances_build: instance %d has a singular or non-finite transform", i);
    const double inv = 1.0 / det;
    const double M[9] = {c00 * inv, (A[2] * A[7] - A[1] * A[8]) * inv, (A[1] * A[5] - A[2] * A[4]) * inv,
                         c01 * inv, (A[0] * A[8] - A[2] * A[6]) * inv, (A[2] * A[3] - A[0] * A[5]) * inv,
                         c02 * inv, (A[1] * A[6] - A[0] * A[7]) * inv, (A[0] * A[4] - A[1] * A[3]) * inv};
    InstRec& r = recs[i];
    std::memset(&r, 0, sizeof(r));
    for (int row = 0; row < 3; ++row) {
      for (int col = 0; col < 3; ++col) r.w2o[4 * row + col] = M[3 * row + col];
      r.w2o[4 * row + 3] = -(M[3 * row] * t[0] + M[3 * row + 1] * t[1] + M[3 * row + 2] * t[2]);
    }
    r.blas = j;
    r.tri_base = int32_t(h_tri_base[i]);
    // world box: the eight corners of the tree's root box, transformed
    const BvhHeader& h = bhdr[j];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int corner = 0; corner < 8; ++corner) {
      const double p[3] = {(corner & 1) ? h.root_hi[0] : h.root_lo[0], (corner & 2) ? h.root_hi[1] : h.root_lo[1],
                           (corner & 4) ? h.root_hi[2] : h.root_lo[2]};
      for (int a = 0; a < 3; ++a) {
        const double w = A[3 * a] * p[0] + A[3 * a + 1] * p[1] + A[3 * a + 2] * p[2] + t[a];
        lo[a] = std::min(lo[a], w);
        hi[a] = std::max(hi[a], w);
      }
    }
    // one primitive whose box is the instance's: vertices lo, hi, (lo.x, hi.y, lo.z)
    double* q = pos.data() + 9 * size_t(i);
    q[0] = lo[0]; q[1] = lo[1]; q[2] = lo[2];
    q[3] = hi[0]; q[4] = hi[1]; q[5] = hi[2];
    q[6] = lo[0]; q[7] = hi[1]; q[8] = lo[2];
  }
  trace_mesh boxes;
  boxes.ntri = n_inst;
  HIP_TRY(ctx, boxes.pos.alloc(pos.size()));
  HIP_TRY(ctx, hipMemcpyAsync(boxes.pos.p, pos.data(), pos.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  trace_bvh* b = new (std::nothrow) trace_bvh();
  if (!b) TRACE_FAIL(ctx, TRACE_E_OOM, "trace_instances_build: host allocation failed");
  int rc = lbvh::build(ctx, &boxes, nullptr, TRACE_KEY_REF63, b);
  if (rc == TRACE_OK) {
    b->instanced = true;
    b->n_inst = n_inst;
    b->blas_depth = depth;
    b->flat_prims = flat;
    hipError_t e = b->inst.alloc(recs.size() * sizeof(InstRec));
    if (e == hipSuccess) e = b->blas.alloc(descs.size() * sizeof(BlasDesc));
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->inst.p, recs.data(), recs.size() * sizeof(InstRec), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->blas.p, descs.data(), descs.size() * sizeof(BlasDesc), hipMemcpyHostToDevice,
                         ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      ctx->err = std::string("trace_instances_build: ") + hipGetErrorString(e);
      rc = e == hipErrorOutOfMemory ? TRACE_E_OOM : TRACE_E_HIP;
    } else {
      rc = ctx_check_device_error(ctx, "trace_instances_build");
    }
  }
  if (rc != TRACE_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return TRACE_OK;
}

int trace_bvh_device_bytes(trace_ctx* ctx, const trace_bvh* bvh, int64_t* bytes) {
  if (!ctx) return TRACE_E_INVALID;
  if (!bvh || !bytes) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_device_bytes: null argument");
  if (bvh->base) {
    *bytes = 0;  // a view owns nothing
    return TRACE_OK;
  }
  const trace_bvh& b = *bvh;
  *bytes = int64_t(b.keys.n * sizeof(uint64_t) + b.perm.n * sizeof(int32_t) + b.tris.n * sizeof(float) +
                   b.prim_box.n * sizeof(float) + b.nodes.n * sizeof(BvhNode) + b.wnodes.n * sizeof(WideNode) +
                   b.parent.n * sizeof(int32_t) + b.run_start.n * sizeof(uint32_t) +
                   b.inner_range.n * sizeof(int32_t) + b.header.n * sizeof(BvhHeader) + b.phr_out.n +
                   b.phr_src.n * sizeof(int32_t) + b.seg_headers.n * sizeof(BvhHeader) + b.virt.n + b.inst.n +
                   b.blas.n);
  return TRACE_OK;
}

}  // extern "C"
