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
//
// Binary leaves of more than kSplitAbove triangles (PHR's makeLeaf flattens
// whole LBVH subtrees into one leaf, thousands of triangles on regular
// geometry such as a heightfield) get a 4-ary sub-tree over their index
// range: level l holds 4^l nodes, node j covering [n j / 4^l, n (j+1) / 4^l)
// of the leaf's n triangles, down to ranges of at most kSplitLeaf.  The
// triangles of a leaf are in Morton order, so index quarters are spatially
// compact.  Only the traversal tree changes: the binary tree, its cost and
// every closest hit stay the same (hits are ordered by t, then by triangle
// position, whatever order nodes are visited in).  Sub-tree nodes follow the
// binary-derived ones; an exclusive scan over per-entry node counts places
// them.  The sub-tree depth is capped so that the traversal stack (three
// entries per level) still holds; every leaf then keeps at most 2048
// triangles, within the 16-bit count of a wide child.
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

// Depth parity by pointer jumping.  jump[i] packs (ancestor << 32 | steps):
// some ancestor of inner node i (-1 once past the root) and how many parent
// links lead there.  Every value ever stored is such a pair, so a thread may
// read another node's word at any moment (one 64-bit load: never torn) and
// jump by it; each jump moves at least one level up.  A thread publishes its
// node's word once, when it has reached the root (the stores write through
// to memory): a node whose ancestor has finished then ends in one more load,
// and the upper levels, whose chains are short, finish first.
__device__ inline uint64_t jump_word(int32_t anc, uint32_t steps) {
  return (uint64_t(uint32_t(anc)) << 32) | steps;
}

// SHIFT 1: parent holds encoded links, (parent << 1) | slot (-1 stays -1
// under the shift).
template <int SHIFT>
__global__ void jump_init_kernel(const int32_t* __restrict__ parent, const BvhHeader* __restrict__ hdr,
                                 int64_t cap, uint64_t* __restrict__ jump) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= cap || i >= hdr->n_inner) return;
  const int32_t p = parent[i] >> SHIFT;
  jump[i] = p >= 0 ? jump_word(p, 1u) : jump_word(-1, 0u);
}

// flag[i] = 1 for binary inner nodes at even depth.
__global__ void flag_kernel(uint64_t* jump, const BvhHeader* __restrict__ hdr, int64_t cap,
                            uint32_t* __restrict__ flag, uint32_t* d_err) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= cap) return;
  const int32_t n_inner = hdr->n_inner;
  if (i >= n_inner) {
    flag[i] = 0u;
    return;
  }
  uint64_t w = jump[i];
  int32_t anc = int32_t(uint32_t(w >> 32));
  uint32_t steps = uint32_t(w);
  for (int it = 0; anc >= 0; ++it) {
    if (it > 4096 || anc >= n_inner || steps > 4096u) {  // deeper than any tree built here: a cycle
      atomicOr(d_err, TRACE_ERR_TREE);
      break;
    }
    const uint64_t v = __hip_atomic_load(jump + anc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    anc = int32_t(uint32_t(v >> 32));
    steps += uint32_t(v);
  }
  __hip_atomic_store(jump + i, jump_word(-1, steps), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  flag[i] = (steps & 1u) == 0u ? 1u : 0u;
}

constexpr int64_t kSplitAbove = 4;  // leaves of more triangles get a sub-tree
constexpr int64_t kSplitLeaf = 4;   // largest leaf of a sub-tree (depth cap aside)

// Levels of the sub-tree of an n-triangle leaf: the fewest h with
// ceil(n / 4^h) <= kSplitLeaf, at most max_levels; 0 for a leaf kept whole.
__device__ inline int split_levels(int64_t n, int max_levels) {
  if (n <= kSplitAbove) return 0;
  int h = 0;
  for (int64_t cap = kSplitLeaf; cap < n && h < max_levels; cap *= 4) ++h;
  return h;
}

// Nodes on the levels above l of a complete 4-ary tree: (4^l - 1) / 3.
__device__ inline uint32_t level_base(int l) { return uint32_t(((uint64_t(1) << (2 * l)) - 1u) / 3u); }

__device__ inline uint32_t split_nodes(int64_t n, int max_levels) {
  return level_base(split_levels(n, max_levels));
}

// Where the sub-trees go: wide index = *base + soff[entry] + local index.
// Entry j < cap: the leaf children of binary node j (slot 0's sub-tree
// first); entry cap + t: the root leaf of tree t.
struct SplitMap {
  const uint32_t* soff;
  const uint32_t* base;
  int max_levels;
};

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
                                  int32_t n_inner, int32_t j, int d, const SplitMap& sm, bool& ok) {
  Slot s;
  const float* box = reinterpret_cast<const float*>(nodes + j) + 6 * d;
  for (int a = 0; a < 3; ++a) {
    s.lo[a] = box[a];
    s.hi[a] = box[3 + a];
  }
  s.used = true;
  const int32_t ref = nodes[j].ref[d], cnt = nodes[j].cnt[d];
  if (cnt > 0) {
    if (cnt <= kSplitAbove) {
      s.ref = ref;
      s.cnt = uint32_t(cnt);
    } else {
      const int32_t c0 = nodes[j].cnt[0];
      const uint32_t skip = d == 1 && c0 > 0 ? split_nodes(c0, sm.max_levels) : 0u;
      s.ref = int32_t(*sm.base + sm.soff[j] + skip);
      s.cnt = 0u;
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
// Quantises four child slots into wide node *dst.
__device__ __attribute__((always_inline)) inline void write_node(const Slot (&s)[4], WideNode* __restrict__ dst,
                                                                  uint32_t* d_err) {
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
  *dst = w;
}

// inv[widx[i]] = i for the binary nodes that become wide nodes.
__global__ void compact_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ widx, int64_t cap,
                               const BvhHeader* __restrict__ hdr, uint32_t* __restrict__ inv) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < cap && i < hdr->n_inner && flag[i]) inv[widx[i]] = uint32_t(i);
}

// One thread per wide node derived from a binary node (w < *n_derived).
__global__ void widen_kernel(const BvhNode* __restrict__ nodes, const BvhHeader* __restrict__ hdr,
                             const uint32_t* __restrict__ inv, const uint32_t* __restrict__ n_derived,
                             const uint32_t* __restrict__ widx, const SplitMap sm, WideNode* __restrict__ out,
                             uint32_t* d_err) {
  const int64_t w = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (w >= *n_derived) return;
  const int32_t i = int32_t(inv[w]);
  const int32_t n_inner = hdr->n_inner;
  if (i < 0 || i >= n_inner) {
    atomicOr(d_err, TRACE_ERR_TREE);
    return;
  }
  bool ok = true;
  Slot s[4];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int32_t cref = nodes[i].ref[c], ccnt = nodes[i].cnt[c];
    if (ccnt > 0) {
      s[2 * c] = child_slot(nodes, widx, n_inner, i, c, sm, ok);
      s[2 * c + 1] = empty_slot();
    } else if (cref >= 0 && cref < n_inner) {
      s[2 * c] = child_slot(nodes, widx, n_inner, cref, 0, sm, ok);
      s[2 * c + 1] = child_slot(nodes, widx, n_inner, cref, 1, sm, ok);
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
  write_node(s, out + w, d_err);
}

// need[e]: sub-tree nodes of entry e (see SplitMap).
__global__ void split_count_kernel(const BvhNode* __restrict__ nodes, const BvhHeader* __restrict__ hdr,
                                   int64_t cap, const BvhHeader* __restrict__ roots, int32_t n_roots,
                                   int max_levels, uint32_t* __restrict__ need) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= cap + n_roots) return;
  uint32_t w = 0u;
  if (e < cap) {
    if (e < hdr->n_inner) {
      const int32_t c0 = nodes[e].cnt[0], c1 = nodes[e].cnt[1];
      if (c0 > 0) w += split_nodes(c0, max_levels);
      if (c1 > 0) w += split_nodes(c1, max_levels);
    }
  } else {
    const int32_t rc = roots[e - cap].root_cnt;
    if (rc > 0) w = split_nodes(rc, max_levels);
  }
  need[e] = w;
}

// need[] from the LBVH leaves: zero for every inner node except the parents of
// leaves over kSplitAbove triangles (rare: a run of equal keys), which add
// their sub-tree sizes; root entries as split_count_kernel.  Reads one run
// start per leaf instead of every node.
__global__ void leaf_need_kernel(const uint32_t* __restrict__ run_start, const int32_t* __restrict__ leaf_links,
                                 const BvhHeader* __restrict__ hdr, int64_t cap, const BvhHeader* __restrict__ roots,
                                 int32_t n_roots, int max_levels, uint32_t* __restrict__ need) {
  const int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (r < n_roots) {
    const int32_t rc = roots[r].root_cnt;
    need[cap + r] = rc > 0 ? split_nodes(rc, max_levels) : 0u;
  }
  if (r >= hdr->n_leaves) return;
  const int64_t c = int64_t(run_start[r + 1]) - int64_t(run_start[r]);
  const int32_t link = leaf_links[r];
  if (c > kSplitAbove && link >= 0 && (link >> 1) < cap) atomicAdd(&need[link >> 1], split_nodes(c, max_levels));
}

// One sub-tree node per thread (grid-stride over the device-side total):
// finds its leaf by binary search over the offsets, then its level and
// index, and writes its four children (sub-tree nodes, or leaves of the
// last level), their boxes the union of the triangle boxes in range.
__global__ void split_kernel(const float* __restrict__ prim_box, int64_t n_prims,
                             const BvhNode* __restrict__ nodes, int64_t cap,
                             const BvhHeader* __restrict__ roots, int32_t n_roots, const SplitMap sm,
                             const uint32_t* __restrict__ n_split, WideNode* __restrict__ out,
                             int64_t out_cap, uint32_t* d_err) {
  const uint32_t total = *n_split, base = *sm.base;
  const int64_t n_entries = cap + n_roots;
  if (int64_t(base) + total > out_cap) {  // leaves overlap: not a tree
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(d_err, TRACE_ERR_TREE);
    return;
  }
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < total; u += gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = n_entries;  // last entry e with soff[e] <= u
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (sm.soff[mid] <= u) lo = mid;
      else hi = mid;
    }
    const int64_t e = lo;
    uint32_t local = u - sm.soff[e];
    uint32_t root = base + sm.soff[e];
    int64_t first, n;
    if (e < cap) {
      const int32_t c0 = nodes[e].cnt[0];
      const uint32_t w0 = c0 > 0 ? split_nodes(c0, sm.max_levels) : 0u;
      const int d = local < w0 ? 0 : 1;
      if (d == 1) {
        local -= w0;
        root += w0;
      }
      first = nodes[e].ref[d];
      n = nodes[e].cnt[d];
    } else {
      first = roots[e - cap].root_ref;
      n = roots[e - cap].root_cnt;
    }
    if (first < 0 || first + n > n_prims) {
      atomicOr(d_err, TRACE_ERR_TREE);
      continue;
    }
    const int levels = split_levels(n, sm.max_levels);
    int l = 0;
    while (l + 1 < levels && level_base(l + 1) <= local) ++l;
    const uint64_t j = local - level_base(l);
    if (levels == 0 || j >= (uint64_t(1) << (2 * l))) {  // offsets disagree with the counts
      atomicOr(d_err, TRACE_ERR_TREE);
      continue;
    }
    const int sh = 2 * (l + 1);
    Slot s[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t a = int64_t((uint64_t(n) * (4 * j + k)) >> sh);
      const int64_t b = int64_t((uint64_t(n) * (4 * j + k + 1)) >> sh);
      if (b <= a) {
        s[k] = empty_slot();
        continue;
      }
      Slot c;
      c.used = true;
      for (int q = 0; q < 3; ++q) {
        c.lo[q] = INFINITY;
        c.hi[q] = -INFINITY;
      }
      for (int64_t t = first + a; t < first + b; ++t) {
        const float* pb = prim_box + 6 * t;
        for (int q = 0; q < 3; ++q) {
          c.lo[q] = fminf(c.lo[q], pb[q]);
          c.hi[q] = fmaxf(c.hi[q], pb[3 + q]);
        }
      }
      if (l + 1 < levels) {
        c.ref = int32_t(root + level_base(l + 1) + uint32_t(4 * j + k));
        c.cnt = 0u;
      } else {
        c.ref = int32_t(first + a);
        c.cnt = uint32_t(b - a);
      }
      s[k] = c;
    }
    write_node(s, out + base + u, d_err);
  }
}

// Wide root of each tree header (the forest's own header is not a tree).
// A root leaf kept whole has wroot -1; n_wide bounds every wide index.
__global__ void wroot_kernel(BvhHeader* __restrict__ hdrs, int32_t count,
                             const uint32_t* __restrict__ widx, const SplitMap sm, int64_t cap,
                             const uint32_t* __restrict__ n_split, BvhHeader* __restrict__ forest_hdr) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t n_wide = int32_t(*sm.base + *n_split);
  if (k == 0 && forest_hdr) {
    forest_hdr->wroot = -1;
    forest_hdr->n_wide = n_wide;
  }
  if (k >= count) return;
  BvhHeader& h = hdrs[k];
  if (h.root_cnt > 0)
    h.wroot = h.root_cnt > kSplitAbove ? int32_t(*sm.base + sm.soff[cap + k]) : -1;
  else
    h.wroot = int32_t(widx[h.root_ref]);
  h.n_wide = n_wide;
}

}  // namespace

namespace {

// Scratch layout of widen(): parent links, depth jumps, the wide-to-binary map, wide flags and indices of the
// binary inner nodes, two totals, then per entry (inner nodes and roots) the
// sub-tree sizes of split leaves, their offsets and the scan's partials.
struct WidePlan {
  size_t o_par, o_jump, o_inv, o_flag, o_idx, o_tot, o_need, o_soff, o_scan, total;
  WidePlan(int64_t cap, int32_t n_roots) {
    const int64_t n_entries = cap + n_roots;
    ScratchPlan plan;
    o_par = plan.take(cap * sizeof(int32_t));
    o_jump = plan.take(cap * sizeof(uint64_t));
    o_inv = plan.take(cap * sizeof(uint32_t));
    o_flag = plan.take(cap * sizeof(uint32_t));
    o_idx = plan.take(cap * sizeof(uint32_t));
    o_tot = plan.take(2 * sizeof(uint32_t));
    o_need = plan.take(n_entries * sizeof(uint32_t));
    o_soff = plan.take(n_entries * sizeof(uint32_t));
    o_scan = plan.take(scan::scratch_words(n_entries) * sizeof(uint32_t));
    total = plan.off;
  }
};

}  // namespace

size_t scratch_bytes(int64_t inner_cap, int32_t n_roots) {
  return WidePlan(std::max<int64_t>(1, inner_cap), std::max<int32_t>(1, n_roots)).total;
}

// A sub-tree of h levels has (4^h - 1) / 3 < 4 n / (3 kSplitLeaf) nodes
// (h is the fewest levels with ceil(n / 4^h) <= kSplitLeaf, so
// n > kSplitLeaf 4^(h-1)), and split leaves hold at most n_prims triangles.
int64_t wide_nodes(int64_t inner_cap, int64_t n_prims) {
  return std::max<int64_t>(1, inner_cap) + 4 * n_prims / (3 * kSplitLeaf) + 1;
}

int widen(trace_ctx* ctx, trace_bvh* bvh, const int32_t* links, const Workspace* ws, const int32_t* leaf_links) {
  const int64_t cap = std::max<int64_t>(1, int64_t(bvh->nodes.n));
  const int32_t n_roots = bvh->nseg > 1 ? bvh->nseg : 1;
  BvhHeader* roots = bvh->nseg > 1 ? bvh->seg_headers.p : bvh->header.p;
  const int64_t n_entries = cap + n_roots;
  const int64_t split_cap = wide_nodes(cap, bvh->n_prims) - cap;
  if (cap + split_cap > int64_t(INT32_MAX))
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "widen: %lld wide nodes exceed the 32-bit index",
               static_cast<long long>(cap + split_cap));
  // Stack bound: a lane holds at most 3 entries per wide level above it.
  // Binary depth D puts binary-derived wide nodes at levels <= (D - 1) / 2,
  // and the kernels keep 128 entries for D <= 63, 512 up to 255.
  const int depth = bvh->depth_bound();
  const int stack = depth > 63 ? 512 : 128;
  const int max_levels = std::min(15, std::max(10, stack / 3 - 1 - std::max(0, (depth - 1) / 2)));
  hipStream_t st = ctx->stream;
  TRACE_ENSURE(ctx, ws, bvh->wnodes, size_t(cap + split_cap));
  const WidePlan plan(cap, n_roots);
  char* base = nullptr;
  int rc = workspace_base(ctx, ws, plan.total, &base, "widen");
  if (rc) return rc;
  auto* parent = reinterpret_cast<int32_t*>(base + plan.o_par);
  auto* flag = reinterpret_cast<uint32_t*>(base + plan.o_flag);
  auto* widx = reinterpret_cast<uint32_t*>(base + plan.o_idx);
  auto* total = reinterpret_cast<uint32_t*>(base + plan.o_tot);  // [0] binary-derived, [1] sub-tree nodes
  auto* need = reinterpret_cast<uint32_t*>(base + plan.o_need);
  auto* soff = reinterpret_cast<uint32_t*>(base + plan.o_soff);
  auto* sscr = reinterpret_cast<uint32_t*>(base + plan.o_scan);
  const int grid = int(ceil_div(cap, 256));
  auto* jump = reinterpret_cast<uint64_t*>(base + plan.o_jump);
  if (links) {
    jump_init_kernel<1><<<grid, 256, 0, st>>>(links, bvh->header.p, cap, jump);
  } else {
    fill_async<int32_t>(st, parent, size_t(cap), -1);
    parent_kernel<<<grid, 256, 0, st>>>(bvh->nodes.p, bvh->header.p, parent);
    jump_init_kernel<0><<<grid, 256, 0, st>>>(parent, bvh->header.p, cap, jump);
  }
  flag_kernel<<<grid, 256, 0, st>>>(jump, bvh->header.p, cap, flag, ctx->d_err);
  scan::exclusive_u32(st, flag, widx, cap, total, sscr);
  if (links && leaf_links) {
    fill_async<uint32_t>(st, need, size_t(cap), 0u);
    leaf_need_kernel<<<int(ceil_div(std::max<int64_t>(bvh->n_prims, n_roots), 256)), 256, 0, st>>>(
        bvh->run_start.p, leaf_links, bvh->header.p, cap, roots, n_roots, max_levels, need);
  } else {
    split_count_kernel<<<int(ceil_div(n_entries, 256)), 256, 0, st>>>(bvh->nodes.p, bvh->header.p, cap, roots,
                                                                      n_roots, max_levels, need);
  }
  scan::exclusive_u32(st, need, soff, n_entries, total + 1, sscr);
  const SplitMap sm{soff, total, max_levels};
  auto* inv = reinterpret_cast<uint32_t*>(base + plan.o_inv);
  compact_kernel<<<grid, 256, 0, st>>>(flag, widx, cap, bvh->header.p, inv);
  widen_kernel<<<grid, 256, 0, st>>>(bvh->nodes.p, bvh->header.p, inv, total, widx, sm, bvh->wnodes.p, ctx->d_err);
  const int sgrid = int(std::min<int64_t>(ceil_div(split_cap, 256), int64_t(ctx->n_cu) * 16));
  split_kernel<<<std::max(1, sgrid), 256, 0, st>>>(bvh->prim_box.p, bvh->n_prims, bvh->nodes.p, cap, roots,
                                                   n_roots, sm, total + 1, bvh->wnodes.p, cap + split_cap,
                                                   ctx->d_err);
  if (bvh->nseg > 1)
    wroot_kernel<<<int(ceil_div(bvh->nseg, 256)), 256, 0, st>>>(bvh->seg_headers.p, bvh->nseg, widx, sm, cap,
                                                               total + 1, bvh->header.p);
  else
    wroot_kernel<<<1, 64, 0, st>>>(bvh->header.p, 1, widx, sm, cap, total + 1, nullptr);
  HIP_TRY(ctx, hipGetLastError());
  return TRACE_OK;
}

}  // namespace wide
