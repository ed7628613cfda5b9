// BVH.traversalSteps (bvh.go:21-44) and HeatMapRenderer.RenderToBuffer
// (render.go:167-197, DefaultTraversalCountShader render.go:250-258) on the
// device.
//
// The count is defined by the reference's own walk, so this kernel follows it
// rather than the wide traversal: an explicit stack that pushes a node's
// children in stored order (the last child is popped first), FP64 slab tests
// (bounding.go:53-91) under the current best t, and FP64 Moeller-Trumbore
// (primitive.go:172-208) in the reference's operation order, over the binary
// tree (LBVH or PHR, whose child order equals the reference's).  An inner node
// whose box the ray enters counts one step.  For FP32-exact inputs the FP32
// boxes and triangles equal the reference's FP64 ones, so every count equals
// the reference's.  One ray per lane: a diagnostic view, not the hot loop.
#include <cfloat>

#include "common.hpp"

namespace steps {
namespace {

struct Ray64 {
  double o[3], d[3], inv[3];
  int sign[3];
};

// camera.go:113-135 newRay
__device__ inline Ray64 make_ray(double ox, double oy, double oz, double dx, double dy, double dz) {
#pragma clang fp contract(off)
  Ray64 r;
  r.o[0] = ox;
  r.o[1] = oy;
  r.o[2] = oz;
  r.d[0] = dx;
  r.d[1] = dy;
  r.d[2] = dz;
  for (int a = 0; a < 3; ++a) {
    r.inv[a] = 1.0 / r.d[a];
    r.sign[a] = r.inv[a] < 0.0 ? 1 : 0;
  }
  return r;
}

// bounding.go:53-91: sign-indexed slab test; a NaN (0 * inf) never rejects.
__device__ inline bool box_hit(const float* lo, const float* hi, const Ray64& r, double tmin, double tmax) {
#pragma clang fp contract(off)
  const double bx0 = r.sign[0] ? hi[0] : lo[0], bx1 = r.sign[0] ? lo[0] : hi[0];
  const double by0 = r.sign[1] ? hi[1] : lo[1], by1 = r.sign[1] ? lo[1] : hi[1];
  const double bz0 = r.sign[2] ? hi[2] : lo[2], bz1 = r.sign[2] ? lo[2] : hi[2];
  double txmin = (bx0 - r.o[0]) * r.inv[0];
  double txmax = (bx1 - r.o[0]) * r.inv[0];
  const double tymin = (by0 - r.o[1]) * r.inv[1];
  const double tymax = (by1 - r.o[1]) * r.inv[1];
  if (txmin > tymax || tymin > txmax) return false;
  if (tymin > txmin) txmin = tymin;
  if (tymax < txmax) txmax = tymax;
  const double tzmin = (bz0 - r.o[2]) * r.inv[2];
  const double tzmax = (bz1 - r.o[2]) * r.inv[2];
  if (txmin > tzmax || tzmin > txmax) return false;
  if (tzmin > txmin) txmin = tzmin;
  if (tzmax < txmax) txmax = tzmax;
  return !(txmax < tmin || txmin > tmax);
}

// primitive.go:172-208 (math.go Dot and Cross, util.go:8-12 ApproxZero):
// tightens t on an accepted hit (t == best accepts, as the reference does).
__device__ inline void tri_hit(const float* __restrict__ p, const Ray64& r, double tmin, double& t) {
#pragma clang fp contract(off)
  const double v0x = p[0], v0y = p[1], v0z = p[2];
  const double e1x = double(p[3]) - v0x, e1y = double(p[4]) - v0y, e1z = double(p[5]) - v0z;
  const double e2x = double(p[6]) - v0x, e2y = double(p[7]) - v0y, e2z = double(p[8]) - v0z;
  const double px = r.d[1] * e2z - r.d[2] * e2y;
  const double py = r.d[2] * e2x - r.d[0] * e2z;
  const double pz = r.d[0] * e2y - r.d[1] * e2x;
  const double det = e1x * px + e1y * py + e1z * pz;
  if (fabs(det) <= 1e-7) return;
  const double inv = 1 / det;
  const double tx = r.o[0] - v0x, ty = r.o[1] - v0y, tz = r.o[2] - v0z;
  const double u = (tx * px + ty * py + tz * pz) * inv;
  if (u < 0 || u > 1) return;
  const double qx = ty * e1z - tz * e1y;
  const double qy = tz * e1x - tx * e1z;
  const double qz = tx * e1y - ty * e1x;
  const double v = (r.d[0] * qx + r.d[1] * qy + r.d[2] * qz) * inv;
  if (v < 0 || u + v > 1) return;
  const double tt = (e2x * qx + e2y * qy + e2z * qz) * inv;
  if (tt < tmin || tt > t) return;
  t = tt;
}

// Sphere.intersected (primitive.go:56-86) in FP64: t must lie strictly
// inside (tmin, t).
__device__ inline void sphere_hit(const float* __restrict__ p, const Ray64& r, double tmin, double& t) {
#pragma clang fp contract(off)
  const double cx = p[0], cy = p[1], cz = p[2], rad = p[3];
  const double ox = r.o[0] - cx, oy = r.o[1] - cy, oz = r.o[2] - cz;
  const double dn = sqrt(r.d[0] * r.d[0] + r.d[1] * r.d[1] + r.d[2] * r.d[2]);
  const double a = dn * dn;
  const double hb = ox * r.d[0] + oy * r.d[1] + oz * r.d[2];
  const double on = sqrt(ox * ox + oy * oy + oz * oz);
  const double c = on * on - rad * rad;
  const double disc = hb * hb - a * c;
  if (disc < 0) return;
  const double sq = sqrt(disc);
  double tt = (-hb - sq) / a;
  if (tt <= tmin || tt >= t) {
    tt = (-hb + sq) / a;
    if (tt <= tmin || tt >= t) return;
  }
  t = tt;
}

// Stack entries: -1 = the root (header), else (binary node << 1) | child slot.
// The reference's stack holds at most one pending sibling per level plus the
// two children just pushed: depth + 2 entries.
// virt (optional): binary nodes that are not nodes of the tree (the binary
// split of an n-ary branch) are walked but not counted.
template <int STACK>
__device__ int32_t walk(const BvhNode* __restrict__ nodes, int64_t n_nodes, const uint8_t* __restrict__ virt,
                        const float* __restrict__ tris, const BvhHeader& hdr, const Ray64& r, double tmin,
                        double tmax, bool& overflow, bool& bad) {
  int32_t st[STACK];
  int sp = 0;
  st[sp++] = -1;
  int32_t count = 0;
  double t = tmax;
  while (sp > 0) {
    const int32_t e = st[--sp];
    float lo[3], hi[3];
    int32_t ref, cnt;
    if (e < 0) {
      for (int a = 0; a < 3; ++a) {
        lo[a] = hdr.root_lo[a];
        hi[a] = hdr.root_hi[a];
      }
      ref = hdr.root_ref;
      cnt = hdr.root_cnt;
    } else {
      const int32_t j = e >> 1, c = e & 1;
      const float* b = reinterpret_cast<const float*>(nodes + j) + 6 * c;
      for (int a = 0; a < 3; ++a) {
        lo[a] = b[a];
        hi[a] = b[3 + a];
      }
      ref = nodes[j].ref[c];
      cnt = nodes[j].cnt[c];
    }
    if (!box_hit(lo, hi, r, tmin, t)) continue;
    if (cnt > 0) {
      if (ref < 0 || cnt > hdr.n_prims - ref) {
        bad = true;
        continue;
      }
      for (int32_t k = ref; k < ref + cnt; ++k) {
        const float* p = tris + 9 * size_t(k);
        if (prim_is_sphere(p)) sphere_hit(p, r, tmin, t);
        else tri_hit(p, r, tmin, t);
      }
    } else {
      if (ref < 0 || ref >= n_nodes || ref >= (1 << 30)) {
        bad = true;
        continue;
      }
      if (sp + 2 > STACK) {
        overflow = true;
        break;
      }
      st[sp++] = ref << 1;
      st[sp++] = (ref << 1) | 1;
      if (!virt || !virt[ref]) ++count;
    }
  }
  return count;
}

template <int STACK>
__global__ __launch_bounds__(256) void steps_kernel(const BvhNode* __restrict__ nodes, int64_t n_nodes,
                                                    const uint8_t* __restrict__ virt,
                                                    const float* __restrict__ tris,
                                                    const BvhHeader* __restrict__ hdr_p,
                                                    const double* __restrict__ rays, int64_t n,
                                                    int32_t* __restrict__ counts, uint32_t* d_err) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const BvhHeader hdr = *hdr_p;
  const double* q = rays + 8 * i;
  const Ray64 r = make_ray(q[0], q[1], q[2], q[4], q[5], q[6]);
  bool overflow = false, bad = false;
  counts[i] = walk<STACK>(nodes, n_nodes, virt, tris, hdr, r, q[3], q[7], overflow, bad);
  if (overflow) atomicOr(d_err, TRACE_ERR_STACK);
  if (bad) atomicOr(d_err, TRACE_ERR_TREE);
}

struct Cam {
  double o[3], ll[3], h[3], v[3];
};

// render.go:167-197: pixel (x, y) -> camera ray through (x / (W-1), y / (H-1))
// (camera.go:95-97: LL + H s + V t - origin), tMin 0.001, tMax MaxFloat64;
// colour by DefaultTraversalCountShader; pixel index y * W + x.
template <int STACK>
__global__ __launch_bounds__(256) void heatmap_kernel(const BvhNode* __restrict__ nodes, int64_t n_nodes,
                                                      const uint8_t* __restrict__ virt,
                                                      const float* __restrict__ tris,
                                                      const BvhHeader* __restrict__ hdr_p, const Cam cam,
                                                      int32_t W, int32_t H, int32_t threshold,
                                                      float* __restrict__ rgb, int32_t* __restrict__ counts,
                                                      uint32_t* d_err) {
#pragma clang fp contract(off)
  const int64_t pix = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (pix >= int64_t(W) * H) return;
  const BvhHeader hdr = *hdr_p;
  const int32_t x = int32_t(pix % W), y = int32_t(pix / W);
  const double s = double(x) / double(W - 1);
  const double t = double(y) / double(H - 1);
  const Ray64 r = make_ray(cam.o[0], cam.o[1], cam.o[2], cam.ll[0] + cam.h[0] * s + cam.v[0] * t - cam.o[0],
                           cam.ll[1] + cam.h[1] * s + cam.v[1] * t - cam.o[1],
                           cam.ll[2] + cam.h[2] * s + cam.v[2] * t - cam.o[2]);
  bool overflow = false, bad = false;
  const int32_t c = walk<STACK>(nodes, n_nodes, virt, tris, hdr, r, 0.001, DBL_MAX, overflow, bad);
  if (counts) counts[pix] = c;
  float* out = rgb + 3 * pix;
  if (c > threshold) {
    out[0] = float(double(c) / double(threshold) * 2);
    out[1] = 0.0f;
    out[2] = 0.0f;
  } else {
    const double f = double(c) / (double(threshold) * 1.25);
    out[0] = 0.0f;
    out[1] = float(f);
    out[2] = float(1 - f);
  }
  if (overflow) atomicOr(d_err, TRACE_ERR_STACK);
  if (bad) atomicOr(d_err, TRACE_ERR_TREE);
}

bool deep(const trace_bvh* bvh) { return bvh->depth_bound() > 63; }

int64_t node_count(const trace_bvh* bvh) {
  return int64_t(bvh->base ? bvh->base->nodes.n : bvh->nodes.n);
}

}  // namespace
}  // namespace steps

extern "C" {

int trace_traversal_steps_device(trace_ctx* ctx, const trace_bvh* bvh, const double* d_rays, int64_t n,
                                 int32_t* d_counts) {
  if (!ctx) return TRACE_E_INVALID;
  if (!bvh || n < 0 || (n > 0 && (!d_rays || !d_counts)))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_traversal_steps: bad arguments");
  if (bvh->instanced) TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_traversal_steps: instanced trees are not walked");
  if (bvh->depth_bound() > 255)
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_traversal_steps: tree depth %d exceeds 255", bvh->depth_bound());
  if (n == 0) return TRACE_OK;
  const int grid = int(ceil_div(n, 256));
  if (steps::deep(bvh))
    steps::steps_kernel<264><<<grid, 256, 0, ctx->stream>>>(bvh->nodes_p(), steps::node_count(bvh), bvh->virt_p(),
                                                            bvh->tris_p(),
                                                            bvh->header_p(), d_rays, n, d_counts, ctx->d_err);
  else
    steps::steps_kernel<72><<<grid, 256, 0, ctx->stream>>>(bvh->nodes_p(), steps::node_count(bvh), bvh->virt_p(),
                                                           bvh->tris_p(),
                                                           bvh->header_p(), d_rays, n, d_counts, ctx->d_err);
  HIP_TRY(ctx, hipGetLastError());
  return TRACE_OK;
}

int trace_traversal_steps(trace_ctx* ctx, const trace_bvh* bvh, const double* h_rays, int64_t n,
                          int32_t* h_counts) {
  if (!ctx) return TRACE_E_INVALID;
  if (!bvh || n < 0 || (n > 0 && (!h_rays || !h_counts)))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_traversal_steps: bad arguments");
  if (n == 0) return TRACE_OK;
  DevArray<char> buf;
  const size_t ray_b = size_t(n) * 8 * sizeof(double);
  HIP_TRY(ctx, buf.alloc(ray_b + size_t(n) * sizeof(int32_t)));
  double* d_rays = reinterpret_cast<double*>(buf.p);
  int32_t* d_counts = reinterpret_cast<int32_t*>(buf.p + ray_b);
  HIP_TRY(ctx, hipMemcpyAsync(d_rays, h_rays, ray_b, hipMemcpyHostToDevice, ctx->stream));
  int rc = trace_traversal_steps_device(ctx, bvh, d_rays, n, d_counts);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(h_counts, d_counts, size_t(n) * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ctx_check_device_error(ctx, "trace_traversal_steps");
}

int trace_render_heatmap_device(trace_ctx* ctx, const trace_bvh* bvh, const trace_camera* cam, int32_t width,
                                int32_t height, int32_t threshold, float* d_rgb, int32_t* d_counts) {
  if (!ctx) return TRACE_E_INVALID;
  if (!bvh || !cam || !d_rgb || width < 2 || height < 2 || threshold < 1)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_heatmap: bad arguments (width, height >= 2, threshold >= 1)");
  if (bvh->instanced) TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_render_heatmap: instanced trees are not walked");
  if (bvh->depth_bound() > 255)
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_render_heatmap: tree depth %d exceeds 255", bvh->depth_bound());
  steps::Cam c;
  for (int a = 0; a < 3; ++a) {
    c.o[a] = cam->origin[a];
    c.ll[a] = cam->lower_left[a];
    c.h[a] = cam->horizontal[a];
    c.v[a] = cam->vertical[a];
  }
  const int grid = int(ceil_div(int64_t(width) * height, 256));
  if (steps::deep(bvh))
    steps::heatmap_kernel<264><<<grid, 256, 0, ctx->stream>>>(bvh->nodes_p(), steps::node_count(bvh),
                                                              bvh->virt_p(), bvh->tris_p(), bvh->header_p(), c, width, height,
                                                              threshold, d_rgb, d_counts, ctx->d_err);
  else
    steps::heatmap_kernel<72><<<grid, 256, 0, ctx->stream>>>(bvh->nodes_p(), steps::node_count(bvh),
                                                             bvh->virt_p(), bvh->tris_p(), bvh->header_p(), c, width, height,
                                                             threshold, d_rgb, d_counts, ctx->d_err);
  HIP_TRY(ctx, hipGetLastError());
  return TRACE_OK;
}

int trace_render_heatmap(trace_ctx* ctx, const trace_bvh* bvh, const trace_camera* cam, int32_t width,
                         int32_t height, int32_t threshold, float* h_rgb, int32_t* h_counts) {
  if (!ctx) return TRACE_E_INVALID;
  if (!h_rgb || width < 2 || height < 2)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_heatmap: bad arguments (width, height >= 2)");
  const size_t npix = size_t(width) * height;
  DevArray<char> buf;
  HIP_TRY(ctx, buf.alloc(npix * (3 * sizeof(float) + sizeof(int32_t))));
  float* d_rgb = reinterpret_cast<float*>(buf.p);
  int32_t* d_counts = reinterpret_cast<int32_t*>(buf.p + npix * 3 * sizeof(float));
  int rc = trace_render_heatmap_device(ctx, bvh, cam, width, height, threshold, d_rgb, d_counts);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(h_rgb, d_rgb, npix * 3 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  if (h_counts)
    HIP_TRY(ctx, hipMemcpyAsync(h_counts, d_counts, npix * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ctx_check_device_error(ctx, "trace_render_heatmap");
}

}  // extern "C"
