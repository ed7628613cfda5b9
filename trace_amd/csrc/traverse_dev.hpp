// Device-side closest-hit traversal shared by the bulk-ray kernel and the
// path tracer.
//
// Semantics follow the reference (bvh.go:179-199, bounding.go:53-91,
// primitive.go:172-208):
//   - slab test against [tmin, best t]; a box is rejected only if it starts
//     beyond the current best hit or ends before tmin;
//   - Moeller-Trumbore with the reference's absolute |det| <= 1e-7 cut and
//     inclusive u, v, t bounds;
//   - on an exact tie in t the reference keeps the triangle its depth-first,
//     children[0]-first walk visits last, i.e. the one with the largest sorted
//     position; the test below accepts an equal t only for a larger position,
//     which reproduces that winner for any visiting order.
// The walk runs over a 4-wide tree (WideNode), nearest child first, with a
// per-lane stack of node indices; leaves are tested as soon as their box is hit.
#pragma once

#include "common.hpp"
#include "traverse.hpp"

namespace traverse {

constexpr float kDetEps = 1e-7f;  // util.go:8 APPROX_THRESH
// Slab exits are stretched by 4 ulp so FP32 rounding never culls a box that
// the exact test would enter (conservative traversal, Ize 2013).
constexpr float kSlabSlack = 1.0f + 4.0f * 1.1920929e-7f;

struct RayPre {
  float ox, oy, oz;
  float ix, iy, iz;  // 1 / direction, IEEE (+-inf on zero components)
  float dx, dy, dz;
  float tmin;
};

__device__ inline RayPre ray_pre(const Ray& r) {
  RayPre p;
  p.ox = r.ox; p.oy = r.oy; p.oz = r.oz;
  p.dx = r.dx; p.dy = r.dy; p.dz = r.dz;
  // hardware reciprocal (1 ulp, +-inf for +-0): box tests stretch their exits
  // by 4 ulp, which covers it
  p.ix = __builtin_amdgcn_rcpf(r.dx);
  p.iy = __builtin_amdgcn_rcpf(r.dy);
  p.iz = __builtin_amdgcn_rcpf(r.dz);
  p.tmin = r.tmin;
  return p;
}

// Entry distance in tn; true if [tn, tf] overlaps [tmin, tmax].
__device__ inline bool slab(const RayPre& r, float lx, float ly, float lz, float hx, float hy,
                            float hz, float tmax, float& tn) {
  const float x0 = (lx - r.ox) * r.ix, x1 = (hx - r.ox) * r.ix;
  const float y0 = (ly - r.oy) * r.iy, y1 = (hy - r.oy) * r.iy;
  const float z0 = (lz - r.oz) * r.iz, z1 = (hz - r.oz) * r.iz;
  const float n = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), r.tmin));
  const float f = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
  tn = n;
  return n <= f * kSlabSlack;
}

// Moeller-Trumbore against sorted triangle k; updates (best_t, best) and the
// barycentrics on acceptance.
__device__ inline void tri_test(const RayPre& r, const float* __restrict__ tris, int32_t k,
                                Hit& h) {
  const float* t = tris + 9 * size_t(k);
  const float v0x = t[0], v0y = t[1], v0z = t[2];
  const float e1x = t[3] - v0x, e1y = t[4] - v0y, e1z = t[5] - v0z;
  const float e2x = t[6] - v0x, e2y = t[7] - v0y, e2z = t[8] - v0z;
  // p = d x e2, det = e1 . p
  const float px = r.dy * e2z - r.dz * e2y;
  const float py = r.dz * e2x - r.dx * e2z;
  const float pz = r.dx * e2y - r.dy * e2x;
  const float det = e1x * px + e1y * py + e1z * pz;
  if (fabsf(det) <= kDetEps) return;
  // hardware reciprocal (1 ulp): this FP32 test only picks the winner, whose
  // t, u, v are re-evaluated in FP64 (refine_hit_fp64)
  const float inv = __builtin_amdgcn_rcpf(det);
  const float tx = r.ox - v0x, ty = r.oy - v0y, tz = r.oz - v0z;
  const float u = (tx * px + ty * py + tz * pz) * inv;
  if (u < 0.0f || u > 1.0f) return;
  // q = tvec x e1
  const float qx = ty * e1z - tz * e1y;
  const float qy = tz * e1x - tx * e1z;
  const float qz = tx * e1y - ty * e1x;
  const float v = (r.dx * qx + r.dy * qy + r.dz * qz) * inv;
  if (v < 0.0f || u + v > 1.0f) return;
  const float tt = (e2x * qx + e2y * qy + e2z * qz) * inv;
  if (tt < r.tmin || tt > h.t) return;
  if (tt == h.t && k < h.pos) return;
  h.t = tt;
  h.t64 = tt;
  h.pos = k;
  h.u = u;
  h.v = v;
  h.front = det > 0.0f;
}

// Sphere.intersected (primitive.go:56-86) in FP32 for sorted primitive k:
// the nearer root if it lies strictly inside (tmin, best t), else the farther
// one (both bounds exclusive, as the reference's).  The discriminant is taken
// as r^2 - |oc - (oc.d / d.d) d|^2 (the same quantity divided by d.d, free of
// the cancellation of the textbook form); the winner is re-solved in FP64.
__device__ inline void sphere_test(const RayPre& r, const float* __restrict__ t, int32_t k, Hit& h) {
  const float ox = r.ox - t[0], oy = r.oy - t[1], oz = r.oz - t[2], rad = t[3];
  const float a = r.dx * r.dx + r.dy * r.dy + r.dz * r.dz;
  const float s = (ox * r.dx + oy * r.dy + oz * r.dz) / a;
  const float lx = ox - s * r.dx, ly = oy - s * r.dy, lz = oz - s * r.dz;
  const float dn = rad * rad - (lx * lx + ly * ly + lz * lz);
  if (dn < 0.0f) return;
  const float q = sqrtf(dn / a);
  float tt = -s - q;
  if (tt <= r.tmin || tt >= h.t) {
    tt = -s + q;
    if (tt <= r.tmin || tt >= h.t) return;
  }
  h.t = tt;
  h.t64 = tt;
  h.pos = k;
  h.u = h.v = 0.0f;
  h.front = true;  // set by refine_hit_fp64
}

// A primitive of a tree that may hold Spheres (SPH): dispatch on the tag.
template <bool SPH>
__device__ inline void prim_test(const RayPre& r, const float* __restrict__ tris, int32_t k, Hit& h) {
  if (SPH && prim_is_sphere(tris + 9 * size_t(k))) sphere_test(r, tris + 9 * size_t(k), k, h);
  else tri_test(r, tris, k, h);
}

// Re-evaluates the accepted triangle in FP64 with the reference's operation
// order (primitive.go:172-208: p = d x e2, det = e1.p, invDet = 1/det,
// t = e2.q * invDet), so the reported t and barycentrics carry FP64 accuracy
// even where the FP32 search was ill-conditioned (grazing rays).
__device__ inline void refine_hit_fp64(const float* __restrict__ tris, const Ray& ray, Hit& h) {
#pragma clang fp contract(off)
  if (h.pos < 0) return;
  const float* t = tris + 9 * size_t(h.pos);
  if (prim_is_sphere(t)) {
    // Sphere.intersected in FP64 (primitive.go:56-86): dirNorm^2, halfB,
    // |oc|^2 - r^2, the nearer root unless it is not beyond tMin; the normal
    // (point - centre) * (1 / r) decides the facing
    const double cx = t[0], cy = t[1], cz = t[2], rad = t[3];
    const double dx = ray.dx, dy = ray.dy, dz = ray.dz;
    const double ox = double(ray.ox) - cx, oy = double(ray.oy) - cy, oz = double(ray.oz) - cz;
    const double dn = sqrt(dx * dx + dy * dy + dz * dz);
    const double a = dn * dn;
    const double hb = ox * dx + oy * dy + oz * dz;
    const double on = sqrt(ox * ox + oy * oy + oz * oz);
    const double c = on * on - rad * rad;
    const double disc = hb * hb - a * c;
    if (disc < 0) return;  // keep the FP32 answer
    const double sq = sqrt(disc);
    double tt = (-hb - sq) / a;
    if (tt <= double(ray.tmin) || tt >= double(ray.tmax)) tt = (-hb + sq) / a;
    const double px = double(ray.ox) + dx * tt, py = double(ray.oy) + dy * tt, pz = double(ray.oz) + dz * tt;
    const double inv = 1 / rad;
    const double nx = (px - cx) * inv, ny = (py - cy) * inv, nz = (pz - cz) * inv;
    h.t64 = tt;
    h.t = float(tt);
    h.front = dx * nx + dy * ny + dz * nz < 0;
    return;
  }
  const double v0x = t[0], v0y = t[1], v0z = t[2];
  const double e1x = double(t[3]) - v0x, e1y = double(t[4]) - v0y, e1z = double(t[5]) - v0z;
  const double e2x = double(t[6]) - v0x, e2y = double(t[7]) - v0y, e2z = double(t[8]) - v0z;
  const double dx = ray.dx, dy = ray.dy, dz = ray.dz;
  const double px = dy * e2z - dz * e2y;
  const double py = dz * e2x - dx * e2z;
  const double pz = dx * e2y - dy * e2x;
  const double det = e1x * px + e1y * py + e1z * pz;
  if (det == 0.0) return;
  const double inv = 1.0 / det;
  const double tx = double(ray.ox) - v0x, ty = double(ray.oy) - v0y, tz = double(ray.oz) - v0z;
  const double qx = ty * e1z - tz * e1y;
  const double qy = tz * e1x - tx * e1z;
  const double qz = tx * e1y - ty * e1x;
  h.u = float((tx * px + ty * py + tz * pz) * inv);
  h.v = float((dx * qx + dy * qy + dz * qz) * inv);
  h.t64 = (e2x * qx + e2y * qy + e2z * qz) * inv;
  h.t = float(h.t64);
}

// Leaf [first, first + cnt) must lie inside the triangle array; a malformed
// reference is skipped and reported instead of read.
template <bool COUNT, bool SPH>
__device__ inline void leaf_test(const RayPre& r, const float* __restrict__ tris, int32_t first,
                                 int32_t cnt, int32_t n_prims, Hit& h, int32_t& n_tris,
                                 bool& bad) {
  if (first < 0 || cnt > n_prims - first) {
    bad = true;
    return;
  }
  for (int32_t k = first; k < first + cnt; ++k) {
    if (COUNT) ++n_tris;
    prim_test<SPH>(r, tris, k, h);
  }
}

// Traversal stack: the top kLdsStack entries of each lane live in LDS, laid
// out [entry][lane] so a wave's pushes and pops hit distinct banks; deeper
// entries (rare: only paths with more than kLdsStack deferred siblings) spill
// to a per-lane private array.
constexpr int kLdsStack = 16;

// The deeper entries' private array is held apart from the stack's pointers:
// an array indexed by a lane's stack depth lives in scratch memory, and a
// struct that embedded it would keep the LDS column pointer and the stride
// there too, so that every push and pop would first load them from scratch
// (and address LDS through flat instructions).
template <int STACK>
struct TravSpill {
  int32_t v[STACK > kLdsStack ? STACK - kLdsStack : 1];
};

// An LDS pointer typed as such: a load that may come from LDS or from the
// private array is then two loads of different pointer types, which the
// compiler cannot fold into one flat load of a selected address.
typedef __attribute__((address_space(3))) int32_t lds_i32;

template <int STACK>
struct TravStack {
  lds_i32* lds;    // this lane's column: entry e at lds[e * stride]
  int stride;
  int32_t* spill;  // TravSpill<STACK>::v: entry e >= kLdsStack at spill[e - kLdsStack]
  __device__ inline void put(int sp, int32_t v) {
    if (sp < kLdsStack) lds[sp * stride] = v;
    else spill[sp - kLdsStack] = v;
  }
  __device__ inline int32_t get(int sp) const {
    int32_t v;
    if (sp < kLdsStack) v = lds[sp * stride];
    else v = spill[sp - kLdsStack];
    return v;
  }
};

// The LDS column of this lane in a __shared__ stack array.
__device__ inline lds_i32* lds_column(int32_t* shared_base) { return (lds_i32*)(shared_base); }

// ---- wide tree (WideNode: 4 children, quantised boxes) -----------------------
//
// Traversal state of one lane, so that a ray can be advanced one node visit at
// a time (the persistent kernel refills lanes whose ray has finished).
struct WideState {
  RayPre r;
  Hit h;
  int32_t node;
  int sp;
  uint32_t visits;
};

// Starts a ray: root box test, root leaf.  Returns true if node visits follow.
template <bool COUNT, bool SPH>
__device__ inline bool wide_begin(const float* __restrict__ tris, const BvhHeader& hdr,
                                  const Ray& ray, WideState& s, int32_t& n_tris, bool& bad) {
  s.r = ray_pre(ray);
  s.h.t = ray.tmax;
  s.h.t64 = ray.tmax;
  s.h.pos = -1;
  s.h.u = s.h.v = 0.0f;
  s.h.front = true;
  s.sp = 0;
  s.visits = 0;
  s.node = hdr.wroot;
  float tn;
  if (!slab(s.r, hdr.root_lo[0], hdr.root_lo[1], hdr.root_lo[2], hdr.root_hi[0], hdr.root_hi[1],
            hdr.root_hi[2], s.h.t, tn))
    return false;
  if (hdr.root_cnt > 0 && hdr.wroot < 0) {  // a root leaf kept whole
    leaf_test<COUNT, SPH>(s.r, tris, hdr.root_ref, hdr.root_cnt, hdr.n_prims, s.h, n_tris, bad);
    return false;
  }
  return true;
}

// One wide-node visit, in three parts:
//   wide_visit   -- all children's boxes are decoded relative to the ray
//                   origin and slab-tested with the near / far planes picked
//                   by the direction's signs; the hit leaf and inner children;
//   leaf tests   -- the primitives of the hit leaves (they can only shorten
//                   the ray, so they run before the inner children are
//                   ordered);
//   wide_advance -- the inner children still in reach are ordered by entry
//                   distance: the nearest is visited next and the others are
//                   pushed farthest first.  Returns false when the ray is done.
// n_bound = the tree's wide-node count, n_prims = its primitive count:
// bounds for every reference followed.
struct NodeVisit {
  float tk[4];      // entry distance per child
  int32_t ref[4];   // child reference: wide node, or first primitive of a leaf
  uint32_t cnt[2];  // leaf sizes, child c in bits 16 (c & 1) .. of cnt[c >> 1]
  uint32_t leaves;  // hit leaf children (bit c)
  uint32_t inner;   // hit inner children (bit c)
  __device__ inline int32_t count(int k) const {
    const uint32_t w = k < 2 ? cnt[0] : cnt[1];
    return int32_t((k & 1) ? w >> 16 : w & 0xffffu);
  }
  __device__ inline int32_t first(int k) const {
    return k == 0 ? ref[0] : k == 1 ? ref[1] : k == 2 ? ref[2] : ref[3];
  }
};

template <bool COUNT>
__device__ inline bool wide_visit(const WideNode* __restrict__ wnodes, uint32_t n_bound, WideState& s,
                                  int32_t& n_nodes, bool& bad, NodeVisit& v) {
  if (uint32_t(s.node) >= n_bound || ++s.visits > n_bound) {
    bad = true;
    return false;
  }
  const uint4* np = reinterpret_cast<const uint4*>(wnodes + s.node);
  const uint4 a = np[0], b = np[1], c = np[2], d = np[3];
  if (COUNT) ++n_nodes;
  const RayPre& r = s.r;
  // plane - origin = fmaf(q, cell, o - origin): exactly (o - origin) for
  // q = 0, and for q > 0 off by far less than the one-cell padding of the
  // stored boxes; times 1/d it keeps the IEEE infinities of axis-parallel rays
  const float rx = __uint_as_float(a.x) - r.ox;
  const float ry = __uint_as_float(a.y) - r.oy;
  const float rz = __uint_as_float(a.z) - r.oz;
  const float sx = __uint_as_float((a.w & 0xffu) << 23);
  const float sy = __uint_as_float(((a.w >> 8) & 0xffu) << 23);
  const float sz = __uint_as_float(((a.w >> 16) & 0xffu) << 23);
  const uint32_t used = a.w >> 24;  // child slots in use
  // near planes are the low ones on axes the ray runs along positively
  const bool px = !(r.ix < 0.0f), py = !(r.iy < 0.0f), pz = !(r.iz < 0.0f);
  const uint32_t nqx = px ? b.x : b.w, fqx = px ? b.w : b.x;
  const uint32_t nqy = py ? b.y : c.x, fqy = py ? c.x : b.y;
  const uint32_t nqz = pz ? b.z : c.y, fqz = pz ? c.y : b.z;
  v.ref[0] = int32_t(c.z);
  v.ref[1] = int32_t(c.w);
  v.ref[2] = int32_t(d.x);
  v.ref[3] = int32_t(d.y);
  v.cnt[0] = d.z;
  v.cnt[1] = d.w;
  const float tlim = s.h.t;
  v.leaves = 0u;
  v.inner = 0u;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int sh = 8 * k;
    const float nx = fmaf(float((nqx >> sh) & 0xffu), sx, rx) * r.ix;
    const float ny = fmaf(float((nqy >> sh) & 0xffu), sy, ry) * r.iy;
    const float nz = fmaf(float((nqz >> sh) & 0xffu), sz, rz) * r.iz;
    const float fx = fmaf(float((fqx >> sh) & 0xffu), sx, rx) * r.ix;
    const float fy = fmaf(float((fqy >> sh) & 0xffu), sy, ry) * r.iy;
    const float fz = fmaf(float((fqz >> sh) & 0xffu), sz, rz) * r.iz;
    const float tn = fmaxf(fmaxf(nx, ny), fmaxf(nz, r.tmin));
    const float tf = fminf(fminf(fx, fy), fminf(fz, tlim));
    v.tk[k] = tn;
    const bool hit = ((used >> k) & 1u) && tn <= tf * kSlabSlack;
    const bool leaf = ((k < 2 ? d.z : d.w) >> (16 * (k & 1))) & 0xffffu;
    if (hit) {
      if (leaf) v.leaves |= 1u << k;
      else v.inner |= 1u << k;
    }
  }
  return true;
}

template <int STACK>
__device__ inline bool wide_advance(WideState& s, TravStack<STACK>& stack, bool& overflow, const NodeVisit& v) {
  // inner children still in reach, sorted by entry distance
  float key[4];
  int32_t id[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const bool live = ((v.inner >> k) & 1u) && v.tk[k] <= s.h.t * kSlabSlack;
    key[k] = live ? v.tk[k] : INFINITY;
    id[k] = live ? v.ref[k] : -1;
  }
#define TRACE_CSWAP(i, j)                      \
  if (key[j] < key[i]) {                       \
    const float tf = key[i];                   \
    key[i] = key[j];                           \
    key[j] = tf;                               \
    const int32_t ti = id[i];                  \
    id[i] = id[j];                             \
    id[j] = ti;                                \
  }
  TRACE_CSWAP(0, 1)
  TRACE_CSWAP(2, 3)
  TRACE_CSWAP(0, 2)
  TRACE_CSWAP(1, 3)
  TRACE_CSWAP(1, 2)
#undef TRACE_CSWAP
  if (id[0] < 0) {
    if (s.sp == 0) return false;
    s.node = stack.get(--s.sp);
    return true;
  }
#pragma unroll
  for (int k = 3; k >= 1; --k) {
    if (id[k] >= 0) {
      if (s.sp < STACK) stack.put(s.sp++, id[k]);
      else overflow = true;
    }
  }
  s.node = id[0];
  return true;
}

// The primitives of a lane's hit leaves back to back in one loop, so that the
// wave iterates as often as its busiest lane has primitives, not once per leaf
// slot times the largest leaf in that slot.  prim(k) tests primitive k.
template <class Prim>
__device__ inline void leaf_prims(const NodeVisit& v, int32_t n_prims, bool& bad, Prim&& prim) {
  uint32_t leaves = v.leaves;
  int32_t cur = 0, end = 0;
  for (;;) {
    if (cur == end && leaves) {
      const int k = __builtin_ctz(leaves);
      leaves &= leaves - 1u;
      const int32_t first = v.first(k), count = v.count(k);
      if (first < 0 || count > n_prims - first) {  // a malformed leaf is reported, not read
        bad = true;
        leaves = 0u;
      } else {
        cur = first;
        end = first + count;
      }
    }
    if (cur == end) break;
    prim(cur);
    ++cur;
  }
}

// A whole visit with the leaf tests in the visiting lane.
template <int STACK, bool COUNT, class Prim>
__device__ inline bool wide_step_with(const WideNode* __restrict__ wnodes, uint32_t n_bound, int32_t n_prims,
                                      WideState& s, TravStack<STACK>& stack, int32_t& n_nodes, bool& overflow,
                                      bool& bad, Prim&& prim) {
  NodeVisit v;
  if (!wide_visit<COUNT>(wnodes, n_bound, s, n_nodes, bad, v)) return false;
  leaf_prims(v, n_prims, bad, prim);
  return wide_advance<STACK>(s, stack, overflow, v);
}

template <int STACK, bool COUNT, bool SPH>
__device__ inline bool wide_step(const WideNode* __restrict__ wnodes,
                                 const float* __restrict__ tris, uint32_t n_bound, int32_t n_prims,
                                 WideState& s, TravStack<STACK>& stack, int32_t& n_nodes,
                                 int32_t& n_tris, bool& overflow, bool& bad) {
  return wide_step_with<STACK, COUNT>(wnodes, n_bound, n_prims, s, stack, n_nodes, overflow, bad,
                                      [&](int32_t k) {
                                        if (COUNT) ++n_tris;
                                        prim_test<SPH>(s.r, tris, k, s.h);
                                      });
}

// Whole-ray wide traversal (used where one lane keeps one ray).
template <int STACK, bool COUNT, bool SPH>
__device__ inline Hit closest_hit_wide(const WideNode* __restrict__ wnodes,
                                       const float* __restrict__ tris, const BvhHeader& hdr,
                                       const Ray& ray, int32_t& n_nodes, int32_t& n_tris,
                                       bool& overflow, bool& bad, int32_t* lds_col, int stride) {
  WideState s;
  TravSpill<STACK> spill;
  TravStack<STACK> stack;
  stack.lds = lds_column(lds_col);
  stack.stride = stride;
  stack.spill = spill.v;
  if (wide_begin<COUNT, SPH>(tris, hdr, ray, s, n_tris, bad))
    while (wide_step<STACK, COUNT, SPH>(wnodes, tris, uint32_t(hdr.n_wide), hdr.n_prims, s, stack,
                                   n_nodes, n_tris, overflow, bad)) {
    }
  return s.h;
}

}  // namespace traverse
