// 4-wide traversal tree with quantised child boxes, derived from a binary tree.
//
// Every binary inner node at even depth (the root has depth 0) becomes a wide
// node; its children are its binary children, except that an inner child (at
// odd depth) is replaced by that child's own two children.  A wide node thus
// has 2 to 4 children, each a leaf or a wide node, and one visit tests what
// would take up to three binary visits.  Depth parity comes from a walk up the
// parent links, wide indices from a scan over binary index order (which keeps
// spatially close subtrees close in memory), and every node is written by one
// thread: no atomics, the same layout on every run.
//
// Child boxes are stored on an 8-bit grid per node and axis (the compressed
// wide BVH of Ylitie, Karras & Laine, "Efficient Incoherent Ray Traversal on
// GPUs Through Compressed Wide BVHs", HPG 2017): origin = the node's FP32 box
// minimum, cell size = a power of two at least 1/253 of the extent and at
// least four FP32 ulps of the coordinates, and each child box widened by one
// cell where the grid allows.  The quantised box is checked with the same
// fmaf decode the traversal uses, so it always contains the child's FP32 box
// and the traversal stays conservative.
#include <algorithm>
#include <cmath>

#include "scan.hpp"
#include "wide.hpp"

namespace wide {
namespace {

__global__ void parent_kernel(const BvhNode* __restrict__ nodes, const BvhHeader* __restrict__ hdr,
                              int32_t* __restrict__ parent) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t n_inner = hdr->n_inner;
  if (i >= n_inner) return;
  const BvhNode b = nodes[i];
  for (int c = 0; c < 2; ++c)
    if (b.cnt[c] == 0 && b.ref[c] >= 0 && b.ref[c] < n_inner) parent[b.ref[c]] = i;
}

// flag[i] = 1 for binary inner nodes at even depth.
__global__ void flag_kernel(const int32_t* __restrict__ parent, const BvhHeader* __restrict__ hdr,
                            int64_t cap, uint32_t* __restrict__ flag, uint32_t* d_err) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= cap) return;
  if (i >= hdr->n_inner) {
    flag[i] = 0u;
    return;
  }
  int d = 0;
  for (int32_t p = parent[i]; p >= 0; p = parent[p]) {
    if (++d > 4096) {  // deeper than any tree built here: a cycle
      atomicOr(d_err, TRACE_ERR_TREE);
      break;
    }
  }
  flag[i] = (d & 1) == 0 ? 1u : 0u;
}

struct Slot {
  float lo[3], hi[3];
  int32_t ref;
  uint32_t cnt;  // 0: inner (ref = wide index)
  bool used;
};

__device__ inline Slot empty_slot() {
  Slot s{};
  s.used = false;
  s.ref = -1;
  s.cnt = 0u;
  return s;
}

// Child d of binary node j (the box is read from memory): a leaf slot, or an
// inner slot pointing at the wide node of that child.
__device__ inline Slot child_slot(const BvhNode* __restrict__ nodes, const uint32_t* __restrict__ widx,
                                  int32_t n_inner, int32_t j, int d, bool& ok) {
  Slot s;
  const float* box = reinterpret_cast<const float*>(nodes + j) + 6 * d;
  for (int a = 0; a < 3; ++a) {
    s.lo[a] = box[a];
    s.hi[a] = box[3 + a];
  }
  s.used = true;
  const int32_t ref = nodes[j].ref[d], cnt = nodes[j].cnt[d];
  if (cnt > 0) {
    if (uint32_t(cnt) < kWideBigLeaf) {
      s.ref = ref;
      s.cnt = uint32_t(cnt);
    } else {
      s.ref = (j << 1) | d;
      s.cnt = kWideBigLeaf;
    }
  } else {
    ok = ok && ref >= 0 && ref < n_inner;
    s.ref = ok ? int32_t(widx[ref]) : -1;
    s.cnt = 0u;
  }
  return s;
}

// e with x < 2^e (x > 0 and normal): the binary exponent of x, plus one
__device__ inline int exp_above(double x) {
  return int((uint64_t(__double_as_longlong(x)) >> 52) & 0x7ff) - 1022;
}

__device__ inline int grid_exponent(float lo, float hi) {
  const double ext = double(hi) - double(lo);
  int k = -126;
  if (ext > 0.0) k = max(k, exp_above(ext / 253.0));  // ext / 253 < 2^k
  const double mag = fmax(fabs(double(lo)), fabs(double(hi)));
  if (mag > 0.0) k = max(k, exp_above(mag) - 22);  // ulp(mag) <= 2^(k - 2)
  return min(k, 127);
}

__device__ inline float decode(uint32_t q, float s, float o) { return fmaf(float(q), s, o); }

// Binary child c of a wide node fills slots 2c and 2c + 1: a leaf child takes
// slot 2c, an inner child (odd depth) is replaced by its two children.
__global__ void widen_kernel(const BvhNode* __restrict__ nodes, const BvhHeader* __restrict__ hdr,
                             const uint32_t* __restrict__ flag, const uint32_t* __restrict__ widx,
                             WideNode* __restrict__ out, uint32_t* d_err) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t n_inner = hdr->n_inner;
  if (i >= n_inner || !flag[i]) return;
  bool ok = true;
  Slot s[4];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int32_t cref = nodes[i].ref[c], ccnt = nodes[i].cnt[c];
    if (ccnt > 0) {
      s[2 * c] = child_slot(nodes, widx, n_inner, i, c, ok);
      s[2 * c + 1] = empty_slot();
    } else if (cref >= 0 && cref < n_inner) {
      s[2 * c] = child_slot(nodes, widx, n_inner, cref, 0, ok);
      s[2 * c + 1] = child_slot(nodes, widx, n_inner, cref, 1, ok);
    } else {
      ok = false;
      s[2 * c] = empty_slot();
      s[2 * c + 1] = empty_slot();
    }
  }
  if (!ok) {
    atomicOr(d_err, TRACE_ERR_TREE);
    return;
  }
  WideNode w;
  uint32_t valid = 0u;
#pragma unroll
  for (int c = 0; c < 4; ++c) valid |= s[c].used ? 1u << c : 0u;
  uint32_t meta = valid << 24;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float lo = INFINITY, hi = -INFINITY;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      lo = s[c].used ? fminf(lo, s[c].lo[a]) : lo;
      hi = s[c].used ? fmaxf(hi, s[c].hi[a]) : hi;
    }
    const int k = grid_exponent(lo, hi);
    const float sc = __uint_as_float(uint32_t(k + 127) << 23);          // 2^k
    const double inv = __longlong_as_double(int64_t(1023 - k) << 52);   // 2^-k
    w.o[a] = lo;
    meta |= uint32_t(k + 127) << (8 * a);
    uint32_t ql = 0, qh = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      uint32_t l = 255u, h = 0u;  // empty slot
      if (s[c].used) {
        const float clo = s[c].lo[a], chi = s[c].hi[a];
        const double fl = floor((double(clo) - double(lo)) * inv) - 1.0;
        const double fh = ceil((double(chi) - double(lo)) * inv) + 1.0;
        l = uint32_t(fmin(fmax(fl, 0.0), 255.0));
        h = uint32_t(fmin(fmax(fh, 0.0), 255.0));
        while (l > 0u && decode(l, sc, lo) > clo) --l;
        while (h < 255u && decode(h, sc, lo) < chi) ++h;
        if (decode(h, sc, lo) < chi) atomicOr(d_err, TRACE_ERR_TREE);  // grid too small
      }
      ql |= l << (8 * c);
      qh |= h << (8 * c);
    }
    w.qlo[a] = ql;
    w.qhi[a] = qh;
  }
  w.meta = meta;
  w.cnt[0] = s[0].cnt | (s[1].cnt << 16);
  w.cnt[1] = s[2].cnt | (s[3].cnt << 16);
#pragma unroll
  for (int c = 0; c < 4; ++c) w.ref[c] = s[c].ref;
  out[widx[i]] = w;
}

// Wide root of each tree header (the forest's own header is not a tree).
__global__ void wroot_kernel(BvhHeader* __restrict__ hdrs, int32_t count,
                             const uint32_t* __restrict__ widx, BvhHeader* __restrict__ forest_hdr) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0 && forest_hdr) forest_hdr->wroot = -1;
  if (k >= count) return;
  BvhHeader& h = hdrs[k];
  h.wroot = h.root_cnt > 0 ? -1 : int32_t(widx[h.root_ref]);
}

}  // namespace

int widen(trace_ctx* ctx, trace_bvh* bvh) {
  const int64_t cap = std::max<int64_t>(1, int64_t(bvh->nodes.n));
  hipStream_t st = ctx->stream;
  HIP_TRY(ctx, bvh->wnodes.alloc(cap));
  ScratchPlan plan;
  const size_t o_par = plan.take(cap * sizeof(int32_t));
  const size_t o_flag = plan.take(cap * sizeof(uint32_t));
  const size_t o_idx = plan.take(cap * sizeof(uint32_t));
  const size_t o_tot = plan.take(sizeof(uint32_t));
  const size_t o_scan = plan.take(scan::scratch_words(cap) * sizeof(uint32_t));
  int rc = ctx_reserve_scratch(ctx, plan.off);
  if (rc) return rc;
  char* base = static_cast<char*>(ctx->scratch);
  auto* parent = reinterpret_cast<int32_t*>(base + o_par);
  auto* flag = reinterpret_cast<uint32_t*>(base + o_flag);
  auto* widx = reinterpret_cast<uint32_t*>(base + o_idx);
  auto* total = reinterpret_cast<uint32_t*>(base + o_tot);
  auto* sscr = reinterpret_cast<uint32_t*>(base + o_scan);
  const int grid = int(ceil_div(cap, 256));
  fill_async<int32_t>(st, parent, size_t(cap), -1);
  parent_kernel<<<grid, 256, 0, st>>>(bvh->nodes.p, bvh->header.p, parent);
  flag_kernel<<<grid, 256, 0, st>>>(parent, bvh->header.p, cap, flag, ctx->d_err);
  scan::exclusive_u32(st, flag, widx, cap, total, sscr);
  widen_kernel<<<grid, 256, 0, st>>>(bvh->nodes.p, bvh->header.p, flag, widx, bvh->wnodes.p, ctx->d_err);
  if (bvh->nseg > 1)
    wroot_kernel<<<int(ceil_div(bvh->nseg, 256)), 256, 0, st>>>(bvh->seg_headers.p, bvh->nseg, widx,
                                                               bvh->header.p);
  else
    wroot_kernel<<<1, 64, 0, st>>>(bvh->header.p, 1, widx, nullptr);
  HIP_TRY(ctx, hipGetLastError());
  return TRACE_OK;
}

}  // namespace wide
