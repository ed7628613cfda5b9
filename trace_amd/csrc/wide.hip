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
};

__device__ inline void leaf_slot(Slot& s, const float* box, int32_t ref, int32_t cnt, int32_t code) {
  for (int a = 0; a < 3; ++a) {
    s.lo[a] = box[a];
    s.hi[a] = box[3 + a];
  }
  if (uint32_t(cnt) < kWideBigLeaf) {
    s.ref = ref;
    s.cnt = uint32_t(cnt);
  } else {
    s.ref = code;
    s.cnt = kWideBigLeaf;
  }
}

// Grid of one axis: cell size 2^k (returned as k) such that the extent spans
// at most 253 cells and a cell is at least four ulps of the coordinates.
__device__ inline int grid_exponent(float lo, float hi) {
  const double ext = double(hi) - double(lo);
  int k = -126;
  if (ext > 0.0) {
    int e;
    frexp(ext / 253.0, &e);  // ext / 253 <= 2^e
    k = max(k, e);
  }
  const double mag = fmax(fabs(double(lo)), fabs(double(hi)));
  if (mag > 0.0) {
    int e;
    frexp(mag, &e);  // mag < 2^e, so ulp(mag) <= 2^(e - 24)
    k = max(k, e - 22);
  }
  return min(k, 127);
}

__device__ inline float decode(uint32_t q, float s, float o) { return fmaf(float(q), s, o); }

__global__ void widen_kernel(const BvhNode* __restrict__ nodes, const BvhHeader* __restrict__ hdr,
                             const uint32_t* __restrict__ flag, const uint32_t* __restrict__ widx,
                             WideNode* __restrict__ out, uint32_t* d_err) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t n_inner = hdr->n_inner;
  if (i >= n_inner || !flag[i]) return;
  const BvhNode b = nodes[i];
  Slot s[4];
  int n = 0;
  for (int c = 0; c < 2; ++c) {
    const float* box = reinterpret_cast<const float*>(&b) + 6 * c;  // child c: lo.xyz, hi.xyz
    if (b.cnt[c] > 0) {
      leaf_slot(s[n++], box, b.ref[c], b.cnt[c], (i << 1) | c);
      continue;
    }
    const int32_t j = b.ref[c];
    if (j < 0 || j >= n_inner) {
      atomicOr(d_err, TRACE_ERR_TREE);
      return;
    }
    const BvhNode cj = nodes[j];
    for (int d = 0; d < 2; ++d) {
      const float* cb = reinterpret_cast<const float*>(&cj) + 6 * d;
      if (cj.cnt[d] > 0) {
        leaf_slot(s[n++], cb, cj.ref[d], cj.cnt[d], (j << 1) | d);
      } else {
        const int32_t g = cj.ref[d];
        if (g < 0 || g >= n_inner) {
          atomicOr(d_err, TRACE_ERR_TREE);
          return;
        }
        Slot& t = s[n++];
        for (int a = 0; a < 3; ++a) {
          t.lo[a] = cb[a];
          t.hi[a] = cb[3 + a];
        }
        t.ref = int32_t(widx[g]);
        t.cnt = 0u;
      }
    }
  }
  WideNode w;
  uint32_t meta = uint32_t(n) << 24;
  for (int a = 0; a < 3; ++a) {
    float lo = s[0].lo[a], hi = s[0].hi[a];
    for (int c = 1; c < n; ++c) {
      lo = fminf(lo, s[c].lo[a]);
      hi = fmaxf(hi, s[c].hi[a]);
    }
    const int k = grid_exponent(lo, hi);
    const float sc = ldexpf(1.0f, k);
    const double inv = ldexp(1.0, -k);
    w.o[a] = lo;
    meta |= uint32_t(k + 127) << (8 * a);
    uint32_t ql = 0, qh = 0;
    for (int c = 0; c < 4; ++c) {
      uint32_t l = 255u, h = 0u;  // empty slot
      if (c < n) {
        const double fl = floor((double(s[c].lo[a]) - double(lo)) * inv) - 1.0;
        const double fh = ceil((double(s[c].hi[a]) - double(lo)) * inv) + 1.0;
        l = uint32_t(fmin(fmax(fl, 0.0), 255.0));
        h = uint32_t(fmin(fmax(fh, 0.0), 255.0));
        while (l > 0u && decode(l, sc, lo) > s[c].lo[a]) --l;
        while (h < 255u && decode(h, sc, lo) < s[c].hi[a]) ++h;
        if (decode(h, sc, lo) < s[c].hi[a]) atomicOr(d_err, TRACE_ERR_TREE);  // grid too small
      }
      ql |= l << (8 * c);
      qh |= h << (8 * c);
    }
    w.qlo[a] = ql;
    w.qhi[a] = qh;
  }
  w.meta = meta;
  w.cnt[0] = w.cnt[1] = 0u;
  for (int c = 0; c < 4; ++c) {
    w.ref[c] = c < n ? s[c].ref : -1;
    if (c < n) w.cnt[c >> 1] |= s[c].cnt << (16 * (c & 1));
  }
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
