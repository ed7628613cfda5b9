// Progressive Hierarchical Refinement (PhrBuilder.BuildFromAuxilary,
// phr.go:47-459) on the GPU: branching factor 2 by the job kernels below,
// 3 to 8 by the group kernel (nary_kernel).
//
// The reference recursion is a tree of independent jobs: a job holds a cut
// (LBVH nodes, a bounding box, a depth), splits it with SweepSAH and refines
// both halves into child jobs.  Here the jobs of one tree level run as one
// launch per job list, and the levels are driven from the device: each
// level's launches read its job counts from the control block, so the host
// enqueues levels in batches and reads the counters once per batch.  Jobs of
// more than 64 cut nodes run one workgroup per job, with the cut staged
// through memory:
//   sort      three stable sorts of the cut by box centre x, y, z (rank by
//             counting: rank = #smaller + #equal-and-earlier), with the z sort
//             fed from the y or x order exactly as phr.go:384-404 does
//   minCost   suffix and prefix box unions and subtree-size sums per order
//             (wave scans), FP64 costs with the reference's operation order,
//             first minimum wins, the no-split cost wins ties
//   refine    keep a node if it is a leaf or SA < S / 2^(alpha*depth+delta),
//             else replace it by its children; one survivor becomes a leaf of
//             all its triangles (phr.go:313-352)
// Smaller jobs (most of the deep levels) keep their cut in registers: 17-64
// nodes one wave per job, 9-16 nodes 16 lanes per job, and at most 8 nodes one
// lane per job with the same steps done sequentially.  Every form performs the
// same FP64 operations in the same order.
// Child jobs, output nodes and leaf lists are appended with atomics; the tree
// *structure* and every triangle range are fixed by the parent, so the result
// is deterministic.  A final pass lays the tree out as the binary traversal
// format, with triangles in PHR leaf order.
//
// The initial cut (phr.go:257-311) is taken in FIFO order from the root with
// the MAX_CUT_SIZE cap -- the order the reference's channel queue follows
// with one worker.  Thresholds per depth are computed on the host with
// std::pow so they equal the oracle's bit for bit.
#include <cmath>
#include <cstring>

#include "onesweep.hpp"
#include "phr.hpp"
#include "scan.hpp"
#include "wide.hpp"

namespace phr {

// Job lists by cut size: one lane per job, 16 lanes per job, one wave per job
// held in registers, and one wave per job staged through memory.
constexpr int kSmallCut = 8;
constexpr int kGroupCut = 16;
constexpr int kWaveCut = 64;
constexpr int kLists = 4;
constexpr int kLarge = 3;
constexpr int kSortLds = 4096;  // cuts of 65 .. this size sort in LDS (bitonic); larger ones by ranking
constexpr int kSmallLds = 1024;  // bitonic size of the device-driven launches without the full sort LDS
constexpr int kMaxDepth = 1024;

struct Flat {  // LBVH as flat ids: 0..U-2 inner nodes, U-1..2U-2 leaves
  float* box;      // 6 per id
  int32_t* child;  // 2 per id, -1 for leaves
  int32_t* size;   // subtree node count (bvh.go:159-167)
  int32_t* prim;   // 2 per id: [first, end) in LBVH sorted order
};

struct Task {
  int32_t off, cnt;  // cut members in the pool
  float box[6];      // cut bounding
  int32_t depth;     // cut depth (initial cut: 1)
  int32_t out;       // output node this job fills
  int32_t prim_off;  // first PHR-ordered triangle of this subtree
  int32_t level;     // depth of the output node
};

// The list a job of k cut nodes goes to.
__device__ inline int route(uint32_t k) {
  return k <= uint32_t(kSmallCut) ? 0 : k <= uint32_t(kGroupCut) ? 1 : k <= uint32_t(kWaveCut) ? 2 : kLarge;
}

struct Lists {  // the next level's job lists
  Task* p[kLists];
  int32_t cap[kLists];
};

// Level counters.  Every job of a level bumps several of them with device-scope
// atomics; each sits on its own 128-byte line so that they do not serialise
// behind one another.
#define TRACE_CTL_FIELD(name) \
  uint32_t name;              \
  uint32_t name##_pad[31]
struct Ctl {
  TRACE_CTL_FIELD(pool);       // pool entries appended for the next level
  TRACE_CTL_FIELD(out);        // output nodes allocated
  TRACE_CTL_FIELD(list);       // leaf-list entries allocated
  TRACE_CTL_FIELD(max_level);
  TRACE_CTL_FIELD(n_leaves);
  TRACE_CTL_FIELD(err);        // 1 capacity, 2 depth table
  TRACE_CTL_FIELD(cut);        // initial cut size
  TRACE_CTL_FIELD(max_m);      // largest cut among the large jobs appended for the next level
  TRACE_CTL_FIELD(groups);     // n-ary build: groups appended for the next step (large list)
  TRACE_CTL_FIELD(groups_small);  // n-ary build: groups of the next step whose cut to split has <= kNaryLaneCut members
  TRACE_CTL_FIELD(done);       // device-driven levels: workgroups of the level's last launch that finished
  unsigned long long work;     // n-ary build: sizes of the cuts handed to SweepSAH (phr.go:221)
  unsigned long long work_pad[15];
  uint32_t jobs[kLists][32];   // jobs appended for the next level, per list (element 0)
  // device-driven levels (advance_level_kernel): the current level's job
  // counts, largest cut, bitonic LDS size and workgroup size of its large
  // jobs, and the levels that held jobs
  uint32_t cur_jobs[kLists];
  uint32_t cur_max_m, cur_lds, sel_b, levels;
  // a rebuild recorded in a graph met a frame beyond its recorded bounds: the
  // layout leaves the tree's arrays as the previous frame left them
  uint32_t stop;
  uint32_t cur_pad[23];
};
#undef TRACE_CTL_FIELD

// ---- branching factors above 2 ---------------------------------------------------
// buildSubTree (phr.go:92-152) keeps splitting the largest of its cuts until it
// holds BranchingFactor of them.  A group is one such job: its cuts, and the
// one to split next.  Every step splits one cut per group (sort, minCost,
// refinement of both halves, as the binary job does), then either hands the
// group on to the next step or, once it holds BranchingFactor cuts or no cut
// has two members, writes its branch: a cut of one member is a leaf, every
// other cut becomes a new group of one cut.  Cuts that are alive at one step
// are disjoint LBVH subtrees, so the members of all of them fit one pool of
// 2U - 1 entries; the members of cuts that wait are carried to the next
// step's pool with the group.
constexpr int kMaxBranching = 8;
constexpr int kNaryLaneCut = 16;  // groups whose cut to split is this small take one lane (nary_lane_kernel)

struct CutRec {
  int32_t off, cnt;        // members in the pool (a PHR-made leaf: cnt 1)
  int32_t depth;           // cut depth (phr.go:360-364)
  int32_t prims;           // triangles under the cut
  int32_t leaf_ls, leaf_lc;  // a PHR-made leaf: its LBVH node list (leaf_lc 0: a cut still to split)
  float box[6];            // cut bounding
};
static_assert(sizeof(CutRec) == 48, "CutRec is 48 bytes");

struct Group {
  int32_t out, level, prim_off, ncuts;
  int32_t split;   // the cut split by the next step
  int32_t pad[3];
  float box[6];    // the job's own cut bounding: its branch's box
  int32_t pad2[2];
  CutRec cuts[kMaxBranching];
};
static_assert(sizeof(Group) % 16 == 0, "Group is a multiple of 16 bytes");

// bounding.go:45-47, FP64 from FP32 box corners (exact differences)
__device__ inline double sa_box(const float* b) {
#pragma clang fp contract(off)
  const double w = double(b[3]) - double(b[0]);
  const double h = double(b[4]) - double(b[1]);
  const double d = double(b[5]) - double(b[2]);
  return 2 * w * h + 2 * w * d + 2 * h * d;
}

__global__ void flat_kernel(const BvhNode* __restrict__ nodes, const int32_t* __restrict__ parent_inner,
                            const int32_t* __restrict__ parent_leaf,
                            const int32_t* __restrict__ inner_range,
                            const uint32_t* __restrict__ run_start, const BvhHeader* __restrict__ hdr,
                            Flat f) {
  const int32_t U = hdr->n_leaves;
  const int32_t ni = U - 1;
  const int32_t id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= 2 * U - 1) return;
  int32_t link;
  if (id < ni) {
    const int32_t lo = inner_range[2 * id], hi = inner_range[2 * id + 1];
    const BvhNode& nd = nodes[id];
    // a leaf child 0 is the leaf at the range's low end, a leaf child 1 the high end
    f.child[2 * id] = nd.cnt[0] > 0 ? ni + lo : nd.ref[0];
    f.child[2 * id + 1] = nd.cnt[1] > 0 ? ni + hi : nd.ref[1];
    f.size[id] = 2 * (hi - lo + 1) - 1;
    f.prim[2 * id] = int32_t(run_start[lo]);
    f.prim[2 * id + 1] = int32_t(run_start[hi + 1]);
    link = parent_inner[id];
  } else {
    const int32_t u = id - ni;
    f.child[2 * id] = -1;
    f.child[2 * id + 1] = -1;
    f.size[id] = 1;
    f.prim[2 * id] = int32_t(run_start[u]);
    f.prim[2 * id + 1] = int32_t(run_start[u + 1]);
    link = parent_leaf[u];
  }
  float* b = f.box + 6 * size_t(id);
  if (link < 0) {
    for (int a = 0; a < 3; ++a) {
      b[a] = hdr->root_lo[a];
      b[3 + a] = hdr->root_hi[a];
    }
  } else {
    const float* src = reinterpret_cast<const float*>(nodes + (link >> 1)) + 6 * (link & 1);
    for (int a = 0; a < 6; ++a) b[a] = src[a];
  }
}

__device__ inline uint32_t block_scan_1024(uint32_t v, uint32_t* s_wave, uint32_t* total) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[w] = x;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int i = 0; i < 16; ++i) {
    if (i < w) off += s_wave[i];
    tot += s_wave[i];
  }
  __syncthreads();
  *total = tot;
  return off + x - v;
}

// phr.go:257-311 in FIFO order (one workgroup of 1024 threads).
__global__ __launch_bounds__(1024) void initial_cut_kernel(Flat f, const double* __restrict__ th, int32_t max_cut,
                                                           int32_t* fa, int32_t* fb, int32_t fcap,
                                                           int32_t* pool, int32_t pool_cap,
                                                           Task* tasks, Ctl* ctl,
                                                           const BvhHeader* __restrict__ hdr) {
  __shared__ uint32_t s_wave[16];
  __shared__ int32_t s_m, s_cut, s_next;
  __shared__ long long s_counter;
  const int tid = threadIdx.x;
  const double th0 = th[0];
  if (tid == 0) {
    fa[0] = 0;  // the LBVH root
    s_m = 1;
    s_cut = 0;
    s_counter = 0;
  }
  __syncthreads();
  int32_t* cur = fa;
  int32_t* nxt = fb;
  while (true) {
    const int32_t m = s_m;
    if (m == 0) break;
    __syncthreads();
    if (tid == 0) s_next = 0;
    __syncthreads();
    for (int32_t c0 = 0; c0 < m; c0 += 1024) {
      const int32_t j = c0 + tid;
      const bool valid = j < m;
      const int32_t id = valid ? cur[j] : 0;
      const bool exp = valid && f.child[2 * id] >= 0 && sa_box(f.box + 6 * size_t(id)) > th0;
      uint32_t etot, ctot, xtot;
      const uint32_t er = block_scan_1024(exp ? 1u : 0u, s_wave, &etot);
      const long long before = s_counter;
      // the counter grows by len(children) - 1 = 1 for every node above the
      // threshold, expanded or not; the node expands while it stays < cap
      const bool expands = exp && (before + er + 1 < max_cut);
      const bool tocut = valid && !expands;
      const uint32_t cr = block_scan_1024(tocut ? 1u : 0u, s_wave, &ctot);
      const uint32_t xr = block_scan_1024(expands ? 1u : 0u, s_wave, &xtot);
      if (tocut) {
        const int32_t p = s_cut + int32_t(cr);
        if (p < pool_cap) pool[p] = id;
        else atomicOr(&ctl->err, 1u);
      }
      if (expands) {
        const int32_t q = s_next + 2 * int32_t(xr);
        if (q + 1 < fcap) {
          nxt[q] = f.child[2 * id];
          nxt[q + 1] = f.child[2 * id + 1];
        } else {
          atomicOr(&ctl->err, 1u);
        }
      }
      __syncthreads();
      if (tid == 0) {
        s_counter += etot;
        s_cut += int32_t(ctot);
        s_next += 2 * int32_t(xtot);
      }
      __syncthreads();
    }
    int32_t* t = cur;
    cur = nxt;
    nxt = t;
    if (tid == 0) s_m = s_next < fcap ? s_next : 0;
    __syncthreads();
  }
  if (tid == 0) {
    Task t;
    t.off = 0;
    t.cnt = s_cut < pool_cap ? s_cut : pool_cap;
    for (int a = 0; a < 3; ++a) {
      t.box[a] = hdr->root_lo[a];
      t.box[3 + a] = hdr->root_hi[a];
    }
    t.depth = 1;
    t.out = 0;
    t.prim_off = 0;
    t.level = 0;
    tasks[0] = t;
    ctl->cut = uint32_t(s_cut);
    ctl->jobs[kLarge][0] = 1;
    ctl->out = 1;  // output node 0 is the root
  }
}

// atomicMax that skips the atomic when the counter already holds v or more
// (most callers of a level find the maximum already there).
__device__ inline void raise_max(uint32_t* c, uint32_t v) {
  if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v) atomicMax(c, v);
}

// ---- wave helpers -------------------------------------------------------------
struct WBox {
  float lo[3], hi[3];
  uint32_t cnt;
};

__device__ inline void wb_empty(WBox& b) {
  for (int a = 0; a < 3; ++a) {
    b.lo[a] = INFINITY;
    b.hi[a] = -INFINITY;
  }
  b.cnt = 0;
}
__device__ inline void wb_merge(WBox& a, const WBox& b) {
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = fminf(a.lo[k], b.lo[k]);
    a.hi[k] = fmaxf(a.hi[k], b.hi[k]);
  }
  a.cnt += b.cnt;
}
__device__ inline WBox wb_shfl_up(const WBox& b, int o) {
  WBox r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = __shfl_up(b.lo[k], o, 64);
    r.hi[k] = __shfl_up(b.hi[k], o, 64);
  }
  r.cnt = __shfl_up(b.cnt, o, 64);
  return r;
}
__device__ inline WBox wb_shfl(const WBox& b, int src) {
  WBox r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = __shfl(b.lo[k], src, 64);
    r.hi[k] = __shfl(b.hi[k], src, 64);
  }
  r.cnt = __shfl(b.cnt, src, 64);
  return r;
}
__device__ inline WBox wb_shfl_xor(const WBox& b, int o) {
  WBox r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = __shfl_xor(b.lo[k], o, 64);
    r.hi[k] = __shfl_xor(b.hi[k], o, 64);
  }
  r.cnt = __shfl_xor(b.cnt, o, 64);
  return r;
}
// inclusive scan across lanes; unions and counts are exact in any order
__device__ inline WBox wb_scan(WBox v) {
  const int lane = threadIdx.x & 63;
  for (int o = 1; o < 64; o <<= 1) {
    const WBox u = wb_shfl_up(v, o);
    if (lane >= o) wb_merge(v, u);
  }
  return v;
}
__device__ inline WBox wb_reduce(WBox v) {
  for (int o = 32; o > 0; o >>= 1) wb_merge(v, wb_shfl_xor(v, o));
  return v;
}
__device__ inline double sa_wb(const WBox& b) {
  const float bb[6] = {b.lo[0], b.lo[1], b.lo[2], b.hi[0], b.hi[1], b.hi[2]};
  return sa_box(bb);
}
__device__ inline uint32_t wave_sum(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline uint32_t wave_excl(uint32_t v, uint32_t* total) {
  const int lane = threadIdx.x & 63;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  *total = __shfl(x, 63, 64);
  return x - v;
}

__device__ inline void load_member(const Flat& f, int32_t node, WBox& b, bool with_size) {
  const float* p = f.box + 6 * size_t(node);
  for (int a = 0; a < 3; ++a) {
    b.lo[a] = p[a];
    b.hi[a] = p[3 + a];
  }
  b.cnt = with_size ? uint32_t(f.size[node]) : 1u;
}

// ---- block helpers: B threads (a multiple of 64) working on one job ----------
// Each helper that crosses waves stages per-wave values in LDS and ends with a
// barrier, so that calls can follow one another.
struct BlockLds {
  WBox box[16];
  uint32_t u[16];
  double d[16];
  int32_t i[16];
};

// Inclusive and exclusive scans of boxes over thread order, and the total.
template <int B>
__device__ inline void block_scan_box(const WBox& x, BlockLds& s, WBox* incl, WBox* excl, WBox* total) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const WBox v = wb_scan(x);
  WBox e = wb_shfl_up(v, 1);
  if (lane == 0) wb_empty(e);
  WBox pre, tot;
  wb_empty(pre);
  if constexpr (B == 64) {
    tot = wb_shfl(v, 63);
  } else {
    if (lane == 63) s.box[w] = v;
    __syncthreads();
    wb_empty(tot);
    for (int i = 0; i < B / 64; ++i) {
      if (i < w) wb_merge(pre, s.box[i]);
      wb_merge(tot, s.box[i]);
    }
    __syncthreads();
  }
  *incl = v;
  wb_merge(*incl, pre);
  *excl = e;
  wb_merge(*excl, pre);
  *total = tot;
}

template <int B>
__device__ inline uint32_t block_excl_u32(uint32_t v, BlockLds& s, uint32_t* total) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  uint32_t wt;
  const uint32_t ex = wave_excl(v, &wt);
  if constexpr (B == 64) {
    *total = wt;
    return ex;
  } else {
    if (lane == 0) s.u[w] = wt;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int i = 0; i < B / 64; ++i) {
      if (i < w) pre += s.u[i];
      tot += s.u[i];
    }
    __syncthreads();
    *total = tot;
    return pre + ex;
  }
}

template <int B>
__device__ inline uint32_t block_sum_u32(uint32_t v, BlockLds& s) {
  uint32_t t;
  block_excl_u32<B>(v, s, &t);
  return t;
}

template <int B>
__device__ inline int32_t block_bcast(int32_t v, BlockLds& s) {
  if constexpr (B == 64) {
    return __shfl(v, 0, 64);
  } else {
    if (threadIdx.x == 0) s.i[0] = v;
    __syncthreads();
    const int32_t r = s.i[0];
    __syncthreads();
    return r;
  }
}

// sort.SliceStable of the cut by box centre on `axis`, starting from order
// `in` (identity when null).  Keys are exact FP64 centres.  Sorting the pairs
// (centre, input position) makes every key distinct, so any correct sort gives
// the stable order: a bitonic network in LDS (P = m rounded up to a power of
// two, padded with +inf) for cuts that fit, rank by counting otherwise; a
// single wave sorts cuts of at most 64 in registers.
template <int B>
__device__ void sort_axis(const Flat& f, const int32_t* ids, const int32_t* in, int32_t* out,
                          int axis, int m, double* keys, double* lkey, int32_t* lpos, int P) {
  const int tid = threadIdx.x;
  if constexpr (B == 64) {
    if (m <= 64) {
      // one pair per lane, bitonic network over lane shuffles
      const int lane = tid;
      double key = INFINITY;
      int32_t pos = lane;
      if (lane < m) {
        const int32_t j = in ? in[lane] : lane;
        const float* b = f.box + 6 * size_t(ids[j]);
        key = (double(b[axis]) + double(b[3 + axis])) * (1.0 / 2.0);
      }
      for (int k = 2; k <= 64; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          const double ok = __shfl_xor(key, j, 64);
          const int32_t op = __shfl_xor(pos, j, 64);
          const bool other_less = ok < key || (ok == key && op < pos);
          const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
          if (keep_min == other_less) {
            key = ok;
            pos = op;
          }
        }
      }
      if (lane < m) out[lane] = in ? in[pos] : pos;
      __syncthreads();
      return;
    }
  }
  if (P > 0) {
    for (int r = tid; r < P; r += B) {
      if (r < m) {
        const int32_t j = in ? in[r] : r;
        const float* b = f.box + 6 * size_t(ids[j]);
        lkey[r] = (double(b[axis]) + double(b[3 + axis])) * (1.0 / 2.0);
        lpos[r] = r;
      } else {
        lkey[r] = INFINITY;
        lpos[r] = 0x7fffffff;
      }
    }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int t = tid; t < (P >> 1); t += B) {
          const int i = 2 * t - (t & (j - 1));  // lower index of pair (i, i + j)
          const int l = i + j;
          const double ka = lkey[i], kb = lkey[l];
          const int32_t pa = lpos[i], pb = lpos[l];
          const bool gt = ka > kb || (ka == kb && pa > pb);
          if (gt == ((i & k) == 0)) {
            lkey[i] = kb;
            lkey[l] = ka;
            lpos[i] = pb;
            lpos[l] = pa;
          }
        }
        __syncthreads();
      }
    }
    for (int r = tid; r < m; r += B) out[r] = in ? in[lpos[r]] : lpos[r];
    __syncthreads();
    return;
  }
  for (int r = tid; r < m; r += B) {
    const int32_t j = in ? in[r] : r;
    const float* b = f.box + 6 * size_t(ids[j]);
    keys[r] = (double(b[axis]) + double(b[3 + axis])) * (1.0 / 2.0);
  }
  __syncthreads();
  for (int r = tid; r < m; r += B) {
    const double kr = keys[r];
    int rank = 0;
    for (int q = 0; q < m; ++q) {
      const double kq = keys[q];
      rank += (kq < kr || (kq == kr && q < r)) ? 1 : 0;
    }
    out[rank] = in ? in[r] : r;
  }
  __syncthreads();
}

struct Split {
  double cost;
  int32_t at;  // 0 = no split
};

// minCost (phr.go:414-459) for the cut in order `ord`: suffix and prefix
// unions and subtree-size sums by block scans over chunks of B positions,
// FP64 costs in the reference's operation order, the first minimum.
template <int B>
__device__ Split min_cost(const Flat& f, const int32_t* ids, const int32_t* ord, int m,
                          const float* cut_box, double* rcost, BlockLds& s) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  WBox carry;
  wb_empty(carry);
  for (int end = m; end > 0; end -= B) {
    const int r = end - 1 - tid;
    WBox v;
    wb_empty(v);
    if (r >= 0) load_member(f, ids[ord[r]], v, true);
    WBox incl, excl, tot;
    block_scan_box<B>(v, s, &incl, &excl, &tot);  // thread t: positions [end - 1 - t, end)
    WBox R = incl;
    wb_merge(R, carry);
    if (r >= 1) rcost[r] = sa_wb(R) * double(R.cnt);
    wb_merge(carry, tot);
  }
  __syncthreads();
  const uint32_t total = carry.cnt;
  wb_empty(carry);
  double best = INFINITY;
  int32_t best_at = m;
  for (int start = 0; start < m; start += B) {
    const int r = start + tid;
    WBox v;
    wb_empty(v);
    if (r < m) load_member(f, ids[ord[r]], v, true);
    WBox incl, excl, tot;
    block_scan_box<B>(v, s, &incl, &excl, &tot);
    WBox P = excl;
    wb_merge(P, carry);
    if (r >= 1 && r < m) {
      const double c = rcost[r] + sa_wb(P) * double(P.cnt);
      if (c < best) {
        best = c;
        best_at = r;
      }
    }
    wb_merge(carry, tot);
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64);
    const int32_t oa = __shfl_xor(best_at, o, 64);
    if (ob < best || (ob == best && oa < best_at)) {
      best = ob;
      best_at = oa;
    }
  }
  if constexpr (B > 64) {
    const int w = tid >> 6;
    if (lane == 0) {
      s.d[w] = best;
      s.i[w] = best_at;
    }
    __syncthreads();
    best = s.d[0];
    best_at = s.i[0];
    for (int i = 1; i < B / 64; ++i) {
      if (s.d[i] < best || (s.d[i] == best && s.i[i] < best_at)) {
        best = s.d[i];
        best_at = s.i[i];
      }
    }
  }
  __syncthreads();
  const double nosplit = sa_box(cut_box) * double(total);
  if (best < nosplit) return Split{best, best_at};
  return Split{nosplit, 0};
}

__device__ inline void write_out(OutNode* out, int32_t o, int kind, int nchild, int32_t c0,
                                 int32_t c1, int32_t list_off, int32_t list_cnt, int32_t prim_off,
                                 int32_t prim_cnt, int32_t level, const float* box) {
  OutNode n;
  n.kind = kind;
  n.nchild = nchild;
  n.child[0] = c0;
  n.child[1] = c1;
  n.list_off = list_off;
  n.list_cnt = list_cnt;
  n.prim_off = prim_off;
  n.prim_cnt = prim_cnt;
  n.level = level;
  for (int a = 0; a < 6; ++a) n.box[a] = box[a];
  n.pad = 0;
  out[o] = n;
}

struct Caps {
  int64_t pool;
  int32_t out, list;
};

// Emits one refined half [a, b) of order C: a leaf when a single node
// survives, otherwise a job of the next level.
template <int B>
__device__ void emit_half(const Flat& f, const int32_t* ids, const int32_t* C, int a, int b,
                          double th, const WBox& hb, int32_t prim_off, int32_t prim_cnt,
                          const Task& t, int32_t o, int32_t* pool_next, const Lists& next, OutNode* out,
                          int32_t* list, Ctl* ctl, const Caps& caps, BlockLds& s) {
  const int tid = threadIdx.x;
  uint32_t k_total = 0;
  for (int r = a + tid; r < b; r += B) {
    const int32_t node = ids[C[r]];
    const bool keep = f.child[2 * node] < 0 || sa_box(f.box + 6 * size_t(node)) < th;
    k_total += keep ? 1u : 2u;
  }
  k_total = block_sum_u32<B>(k_total, s);
  const float box[6] = {hb.lo[0], hb.lo[1], hb.lo[2], hb.hi[0], hb.hi[1], hb.hi[2]};
  if (k_total == 1) {  // one kept node: a leaf of all its triangles
    if (tid == 0) {
      const int32_t ls = int32_t(atomicAdd(&ctl->list, 1u));
      if (ls < caps.list) {
        list[ls] = ids[C[a]];
        write_out(out, o, 1, 0, -1, -1, ls, 1, prim_off, prim_cnt, t.level + 1, box);
        raise_max(&ctl->max_level, uint32_t(t.level + 1));
      } else {
        atomicOr(&ctl->err, 1u);
      }
    }
    return;
  }
  const int dst = route(k_total);
  int32_t slot = 0, poff = 0;
  if (tid == 0) {
    slot = int32_t(atomicAdd(&ctl->jobs[dst][0], 1u));
    poff = int32_t(atomicAdd(&ctl->pool, k_total));
  }
  slot = block_bcast<B>(slot, s);
  poff = block_bcast<B>(poff, s);
  if (slot >= next.cap[dst] || int64_t(poff) + k_total > caps.pool) {
    if (tid == 0) atomicOr(&ctl->err, 1u);
    return;
  }
  uint32_t carry = 0;
  for (int r0 = a; r0 < b; r0 += B) {
    const int r = r0 + tid;
    int32_t node = 0;
    uint32_t k = 0;
    if (r < b) {
      node = ids[C[r]];
      const bool keep = f.child[2 * node] < 0 || sa_box(f.box + 6 * size_t(node)) < th;
      k = keep ? 1u : 2u;
    }
    uint32_t tot;
    const uint32_t ex = block_excl_u32<B>(k, s, &tot);
    const int32_t p = poff + int32_t(carry + ex);
    if (k == 1) {
      pool_next[p] = node;
    } else if (k == 2) {
      pool_next[p] = f.child[2 * node];
      pool_next[p + 1] = f.child[2 * node + 1];
    }
    carry += tot;
  }
  if (tid == 0) {
    if (dst == kLarge) raise_max(&ctl->max_m, k_total);
    Task n;
    n.off = poff;
    n.cnt = int32_t(k_total);
    for (int q = 0; q < 6; ++q) n.box[q] = box[q];
    n.depth = t.depth + 1;
    n.out = o;
    n.prim_off = prim_off;
    n.level = t.level + 1;
    next.p[dst][slot] = n;
  }
}

// One workgroup of B threads per job of the current level (phr.go:92-152
// with BranchingFactor 2).
// One job of the large list, by the whole workgroup (s_dyn: lds_p keys then
// lds_p positions of dynamic LDS, 0 = no LDS sort).
template <int B>
__device__ void task_job(const Task t, Flat f, const double* __restrict__ th, const int32_t* __restrict__ pool,
                         int32_t* __restrict__ pool_next, Lists next, int32_t* __restrict__ ordbuf, double* gkey,
                         double* rcost_buf, OutNode* __restrict__ out, int32_t* __restrict__ list, Ctl* ctl,
                         Caps caps, int lds_p, double* s_dyn, BlockLds& s_blk) {
  const int tid = threadIdx.x;
  const int m = t.cnt;
  // bitonic size of this job: its own cut rounded up to a power of two (the
  // level's LDS holds the largest), or 0 to sort by ranking
  int P = 0;
  if (m <= lds_p) {
    P = 128;
    while (P < m) P <<= 1;
  }
  double* lkey = s_dyn;
  int32_t* lpos = reinterpret_cast<int32_t*>(s_dyn + lds_p);
  const int32_t* ids = pool + t.off;
  int32_t* oX = ordbuf + t.off;
  int32_t* oY = ordbuf + caps.pool + t.off;
  int32_t* oZ = ordbuf + 2 * caps.pool + t.off;
  double* keys = gkey + t.off;
  double* rc = rcost_buf + t.off;
  if (t.depth >= kMaxDepth) {
    if (tid == 0) atomicOr(&ctl->err, 2u);
    return;
  }

  // SweepSAH (phr.go:368-405)
  sort_axis<B>(f, ids, nullptr, oX, 0, m, keys, lkey, lpos, P);
  sort_axis<B>(f, ids, oX, oY, 1, m, keys, lkey, lpos, P);
  const Split x = min_cost<B>(f, ids, oX, m, t.box, rc, s_blk);
  const Split y = min_cost<B>(f, ids, oY, m, t.box, rc, s_blk);
  sort_axis<B>(f, ids, x.cost < y.cost ? oY : oX, oZ, 2, m, keys, lkey, lpos, P);
  const Split z = min_cost<B>(f, ids, oZ, m, t.box, rc, s_blk);
  const int32_t* C;
  Split s;
  if (x.cost < y.cost) {
    if (x.cost < z.cost) { C = oX; s = x; } else { C = oZ; s = z; }
  } else {
    if (y.cost < z.cost) { C = oY; s = y; } else { C = oZ; s = z; }
  }

  if (s.at > 0) {
    WBox L, R;
    wb_empty(L);
    wb_empty(R);
    uint32_t pl = 0, pr = 0;
    for (int r = tid; r < m; r += B) {
      const int32_t node = ids[C[r]];
      WBox v;
      load_member(f, node, v, false);
      const uint32_t pc = uint32_t(f.prim[2 * node + 1] - f.prim[2 * node]);
      if (r < s.at) { wb_merge(L, v); pl += pc; } else { wb_merge(R, v); pr += pc; }
    }
    WBox i_, e_, Lt, Rt;
    block_scan_box<B>(L, s_blk, &i_, &e_, &Lt);
    block_scan_box<B>(R, s_blk, &i_, &e_, &Rt);
    pl = block_sum_u32<B>(pl, s_blk);
    pr = block_sum_u32<B>(pr, s_blk);
    int32_t o = 0;
    if (tid == 0) o = int32_t(atomicAdd(&ctl->out, 2u));
    o = block_bcast<B>(o, s_blk);
    if (o + 1 >= caps.out) {
      if (tid == 0) atomicOr(&ctl->err, 1u);
      return;
    }
    if (tid == 0)
      write_out(out, t.out, 0, 2, o, o + 1, -1, 0, t.prim_off, int32_t(pl + pr), t.level, t.box);
    const double thd = th[t.depth];  // refined() uses the split cut's depth
    emit_half<B>(f, ids, C, 0, s.at, thd, Lt, t.prim_off, int32_t(pl), t, o, pool_next, next, out, list, ctl,
                 caps, s_blk);
    emit_half<B>(f, ids, C, s.at, m, thd, Rt, t.prim_off + int32_t(pl), int32_t(pr), t, o + 1, pool_next, next,
                 out, list, ctl, caps, s_blk);
  } else {
    // no split beats the whole cut: one leaf of every triangle, in order C
    uint32_t pc = 0;
    for (int r = tid; r < m; r += B) {
      const int32_t node = ids[C[r]];
      pc += uint32_t(f.prim[2 * node + 1] - f.prim[2 * node]);
    }
    pc = block_sum_u32<B>(pc, s_blk);
    int32_t o = 0, ls = 0;
    if (tid == 0) {
      o = int32_t(atomicAdd(&ctl->out, 1u));
      ls = int32_t(atomicAdd(&ctl->list, uint32_t(m)));
    }
    o = block_bcast<B>(o, s_blk);
    ls = block_bcast<B>(ls, s_blk);
    if (o >= caps.out || ls + m > caps.list) {
      if (tid == 0) atomicOr(&ctl->err, 1u);
      return;
    }
    for (int r = tid; r < m; r += B) list[ls + r] = ids[C[r]];
    if (tid == 0) {
      write_out(out, o, 1, 0, -1, -1, ls, m, t.prim_off, int32_t(pc), t.level + 1, t.box);
      write_out(out, t.out, 0, 1, o, -1, -1, 0, t.prim_off, int32_t(pc), t.level, t.box);
      raise_max(&ctl->max_level, uint32_t(t.level + 1));
    }
  }
}

// Device-driven levels: the workgroup size advance_level_kernel chose for
// this level's large jobs runs them, grid-stride; lds_cap bounds the bitonic
// size this launch reserved LDS for (larger cuts sort by ranking, which
// gives the same order).
template <int B>
__global__ __launch_bounds__(B) void task_dev_kernel(Flat f, const double* __restrict__ th,
                                                     const Task* __restrict__ tasks,
                                                     const int32_t* __restrict__ pool,
                                                     int32_t* __restrict__ pool_next, Lists next,
                                                     int32_t* __restrict__ ordbuf, double* gkey,
                                                     double* rcost_buf, OutNode* __restrict__ out,
                                                     int32_t* __restrict__ list, Ctl* ctl, Caps caps,
                                                     int lds_cap) {
  extern __shared__ double s_dyn[];
  __shared__ BlockLds s_blk;
  if (ctl->sel_b != uint32_t(B)) return;
  const uint32_t n = ctl->cur_jobs[kLarge];
  const int lds_p = min(int(ctl->cur_lds), lds_cap);
  for (uint32_t j = blockIdx.x; j < n; j += gridDim.x) {
    task_job<B>(tasks[j], f, th, pool, pool_next, next, ordbuf, gkey, rcost_buf, out, list, ctl, caps, lds_p,
                s_dyn, s_blk);
    __syncthreads();  // the next job reuses the LDS
  }
}

// ---- n-ary jobs: one workgroup per group --------------------------------------------
struct NCaps {
  int64_t pool;
  int32_t out, list, groups;
};

// refined() (phr.go:313-338) of one half [a, b) of order C with the split
// cut's threshold: the single survivor becomes a leaf of all its triangles,
// otherwise the refined members are appended to the next pool.  The record
// is the same in every thread.
template <int B>
__device__ CutRec refine_half(const Flat& f, const int32_t* ids, const int32_t* C, int a, int b, double th,
                              const WBox& hb, int32_t prims, int32_t depth, int32_t* pool_next, int32_t* list,
                              Ctl* ctl, const NCaps& caps, BlockLds& s) {
  const int tid = threadIdx.x;
  uint32_t k_total = 0;
  for (int r = a + tid; r < b; r += B) {
    const int32_t node = ids[C[r]];
    const bool keep = f.child[2 * node] < 0 || sa_box(f.box + 6 * size_t(node)) < th;
    k_total += keep ? 1u : 2u;
  }
  k_total = block_sum_u32<B>(k_total, s);
  CutRec rec;
  rec.depth = depth;
  rec.prims = prims;
  const float box[6] = {hb.lo[0], hb.lo[1], hb.lo[2], hb.hi[0], hb.hi[1], hb.hi[2]};
  for (int q = 0; q < 6; ++q) rec.box[q] = box[q];
  if (k_total == 1) {
    int32_t ls = 0;
    if (tid == 0) {
      ls = int32_t(atomicAdd(&ctl->list, 1u));
      if (ls < caps.list) list[ls] = ids[C[a]];
      else atomicOr(&ctl->err, 1u);
    }
    ls = block_bcast<B>(ls, s);
    rec.off = 0;
    rec.cnt = 1;
    rec.leaf_ls = ls;
    rec.leaf_lc = 1;
    return rec;
  }
  int32_t poff = 0;
  if (tid == 0) poff = int32_t(atomicAdd(&ctl->pool, k_total));
  poff = block_bcast<B>(poff, s);
  rec.off = poff;
  rec.cnt = int32_t(k_total);
  rec.leaf_ls = 0;
  rec.leaf_lc = 0;
  if (int64_t(poff) + k_total > caps.pool) {
    if (tid == 0) atomicOr(&ctl->err, 1u);
    rec.cnt = 0;
    return rec;
  }
  uint32_t carry = 0;
  for (int r0 = a; r0 < b; r0 += B) {
    const int r = r0 + tid;
    int32_t node = 0;
    uint32_t k = 0;
    if (r < b) {
      node = ids[C[r]];
      const bool keep = f.child[2 * node] < 0 || sa_box(f.box + 6 * size_t(node)) < th;
      k = keep ? 1u : 2u;
    }
    uint32_t tot;
    const uint32_t ex = block_excl_u32<B>(k, s, &tot);
    const int32_t p = poff + int32_t(carry + ex);
    if (k == 1) {
      pool_next[p] = node;
    } else if (k == 2) {
      pool_next[p] = f.child[2 * node];
      pool_next[p + 1] = f.child[2 * node + 1];
    }
    carry += tot;
  }
  return rec;
}

// The groups of a step sit in one buffer of caps.groups records: those whose
// cut to split has more than kNaryLaneCut members from the front (ctl->groups),
// the others from the back (ctl->groups_small; entry s at caps.groups - 1 - s).
// A group's cuts are disjoint sets of at least two pool entries, so the two
// lists together hold at most caps.groups records.
__device__ inline Group* group_slot(Group* next, const NCaps& caps, Ctl* ctl, uint32_t split_cnt) {
  if (split_cnt <= uint32_t(kNaryLaneCut)) {
    const int32_t s = int32_t(atomicAdd(&ctl->groups_small, 1u));
    if (s >= caps.groups) return nullptr;
    return next + (caps.groups - 1 - s);
  }
  const int32_t s = int32_t(atomicAdd(&ctl->groups, 1u));
  if (s >= caps.groups) return nullptr;
  raise_max(&ctl->max_m, split_cnt);
  return next + s;
}

// One step of every group: one split (phr.go:102-127 for BranchingFactor bf),
// then the group moves on or writes its branch.
template <int B>
__global__ __launch_bounds__(B) void nary_kernel(Flat f, const double* __restrict__ th,
                                                 const Group* __restrict__ groups, Group* __restrict__ groups_next,
                                                 const int32_t* __restrict__ pool, int32_t* __restrict__ pool_next,
                                                 int32_t* __restrict__ ordbuf, double* gkey, double* rcost_buf,
                                                 OutNode* __restrict__ out, int32_t* __restrict__ list, Ctl* ctl,
                                                 NCaps caps, int lds_p, int bf) {
  extern __shared__ double s_dyn[];
  __shared__ BlockLds s_blk;
  __shared__ Group g;
  const int tid = threadIdx.x;
  {
    const int32_t* src = reinterpret_cast<const int32_t*>(groups + blockIdx.x);
    int32_t* dst = reinterpret_cast<int32_t*>(&g);
    for (int w = tid; w < int(sizeof(Group) / 4); w += B) dst[w] = src[w];
  }
  __syncthreads();
  const int si = g.split;
  const CutRec c = g.cuts[si];
  const int m = c.cnt;
  if (c.depth >= kMaxDepth || si < 0 || si >= g.ncuts || m < 2) {
    if (tid == 0) atomicOr(&ctl->err, 2u);
    return;
  }
  int P = 0;
  if (m <= lds_p) {
    P = 128;
    while (P < m) P <<= 1;
  }
  double* lkey = s_dyn;
  int32_t* lpos = reinterpret_cast<int32_t*>(s_dyn + lds_p);
  const int32_t* ids = pool + c.off;
  int32_t* oX = ordbuf + c.off;
  int32_t* oY = ordbuf + caps.pool + c.off;
  int32_t* oZ = ordbuf + 2 * caps.pool + c.off;
  double* keys = gkey + c.off;
  double* rc = rcost_buf + c.off;

  // SweepSAH (phr.go:368-405), as the binary job does it
  sort_axis<B>(f, ids, nullptr, oX, 0, m, keys, lkey, lpos, P);
  sort_axis<B>(f, ids, oX, oY, 1, m, keys, lkey, lpos, P);
  const Split x = min_cost<B>(f, ids, oX, m, c.box, rc, s_blk);
  const Split y = min_cost<B>(f, ids, oY, m, c.box, rc, s_blk);
  sort_axis<B>(f, ids, x.cost < y.cost ? oY : oX, oZ, 2, m, keys, lkey, lpos, P);
  const Split z = min_cost<B>(f, ids, oZ, m, c.box, rc, s_blk);
  const int32_t* C;
  Split sp;
  if (x.cost < y.cost) {
    if (x.cost < z.cost) { C = oX; sp = x; } else { C = oZ; sp = z; }
  } else {
    if (y.cost < z.cost) { C = oY; sp = y; } else { C = oZ; sp = z; }
  }
  if (tid == 0) atomicAdd(&ctl->work, (unsigned long long)m);  // buildSubTreeCost's counter

  CutRec first, second;
  bool two = false;
  if (sp.at > 0) {
    WBox L, R;
    wb_empty(L);
    wb_empty(R);
    uint32_t pl = 0, pr = 0;
    for (int r = tid; r < m; r += B) {
      const int32_t node = ids[C[r]];
      WBox v;
      load_member(f, node, v, false);
      const uint32_t pc = uint32_t(f.prim[2 * node + 1] - f.prim[2 * node]);
      if (r < sp.at) { wb_merge(L, v); pl += pc; } else { wb_merge(R, v); pr += pc; }
    }
    WBox i_, e_, Lt, Rt;
    block_scan_box<B>(L, s_blk, &i_, &e_, &Lt);
    block_scan_box<B>(R, s_blk, &i_, &e_, &Rt);
    pl = block_sum_u32<B>(pl, s_blk);
    pr = block_sum_u32<B>(pr, s_blk);
    const double thd = th[c.depth];  // refined() uses the split cut's depth
    first = refine_half<B>(f, ids, C, 0, sp.at, thd, Lt, int32_t(pl), c.depth + 1, pool_next, list, ctl, caps, s_blk);
    second = refine_half<B>(f, ids, C, sp.at, m, thd, Rt, int32_t(pr), c.depth + 1, pool_next, list, ctl, caps,
                            s_blk);
    two = true;
  } else {
    // no split beats the whole cut: a leaf of every member, in order C (phr.go:121-126)
    int32_t ls = 0;
    if (tid == 0) ls = int32_t(atomicAdd(&ctl->list, uint32_t(m)));
    ls = block_bcast<B>(ls, s_blk);
    if (ls + m > caps.list) {
      if (tid == 0) atomicOr(&ctl->err, 1u);
      return;
    }
    for (int r = tid; r < m; r += B) list[ls + r] = ids[C[r]];
    first = c;
    first.off = 0;
    first.cnt = 1;
    first.leaf_ls = ls;
    first.leaf_lc = m;
  }
  // the group's cuts after this split: the split one replaced, the right half appended
  const int n_old = g.ncuts;
  __syncthreads();
  if (tid == 0) {
    g.cuts[si] = first;
    if (two) g.cuts[g.ncuts++] = second;
  }
  __syncthreads();
  // carry the waiting cuts' members into the next pool
  for (int j = 0; j < n_old; ++j) {
    if (j == si || g.cuts[j].leaf_lc > 0) continue;
    const int32_t cnt = g.cuts[j].cnt;
    int32_t poff = 0;
    if (tid == 0) poff = int32_t(atomicAdd(&ctl->pool, uint32_t(cnt)));
    poff = block_bcast<B>(poff, s_blk);
    if (int64_t(poff) + cnt > caps.pool) {
      if (tid == 0) atomicOr(&ctl->err, 1u);
      return;
    }
    const int32_t* from = pool + g.cuts[j].off;
    for (int r = tid; r < cnt; r += B) pool_next[poff + r] = from[r];
    __syncthreads();
    if (tid == 0) g.cuts[j].off = poff;
    __syncthreads();
  }
  if (tid != 0) return;
  // the largest cut, the first of equal sizes (phr.go:104-111)
  int32_t mx = 0, mi = 0;
  for (int j = 0; j < g.ncuts; ++j)
    if (g.cuts[j].cnt > mx) {
      mx = g.cuts[j].cnt;
      mi = j;
    }
  if (g.ncuts < bf && mx > 1) {
    Group* dst = group_slot(groups_next, caps, ctl, uint32_t(mx));
    if (!dst) {
      atomicOr(&ctl->err, 1u);
      return;
    }
    g.split = mi;
    *dst = g;
    return;
  }
  // the branch (box: the job's cut bounding) and its children in cut order
  const int32_t k = g.ncuts;
  const int32_t o = int32_t(atomicAdd(&ctl->out, uint32_t(k)));
  if (o + k > caps.out) {
    atomicOr(&ctl->err, 1u);
    return;
  }
  int32_t p = g.prim_off;
  for (int j = 0; j < k; ++j) {
    const CutRec& cj = g.cuts[j];
    if (cj.leaf_lc > 0) {
      write_out(out, o + j, 1, 0, -1, -1, cj.leaf_ls, cj.leaf_lc, p, cj.prims, g.level + 1, cj.box);
      raise_max(&ctl->max_level, uint32_t(g.level + 1));
    } else {
      Group* dst = group_slot(groups_next, caps, ctl, uint32_t(cj.cnt));
      if (!dst) {
        atomicOr(&ctl->err, 1u);
        return;
      }
      Group n;
      n.out = o + j;
      n.level = g.level + 1;
      n.prim_off = p;
      n.ncuts = 1;
      n.split = 0;
      for (int q = 0; q < 3; ++q) n.pad[q] = 0;
      for (int q = 0; q < 6; ++q) n.box[q] = cj.box[q];
      n.pad2[0] = n.pad2[1] = 0;
      n.cuts[0] = cj;
      *dst = n;
    }
    p += cj.prims;
  }
  write_out(out, g.out, 0, k, o, k > 1 ? o + 1 : -1, -1, 0, g.prim_off, p - g.prim_off, g.level, g.box);
}

// The first group: the initial cut's job as initial_cut_kernel left it.
__global__ void nary_root_kernel(const Task* __restrict__ t0, int32_t n, Group* __restrict__ g0) {
  const Task t = *t0;
  Group g;
  g.out = t.out;
  g.level = t.level;
  g.prim_off = t.prim_off;
  g.ncuts = 1;
  g.split = 0;
  for (int q = 0; q < 3; ++q) g.pad[q] = 0;
  for (int q = 0; q < 6; ++q) g.box[q] = t.box[q];
  g.pad2[0] = g.pad2[1] = 0;
  CutRec c;
  c.off = t.off;
  c.cnt = t.cnt;
  c.depth = t.depth;
  c.prims = n;
  c.leaf_ls = 0;
  c.leaf_lc = 0;
  for (int q = 0; q < 6; ++q) c.box[q] = t.box[q];
  g.cuts[0] = c;
  *g0 = g;
}

// ---- small jobs: one lane per job ----------------------------------------------
// The deep levels hold hundreds of thousands of jobs whose cuts have a few
// nodes; a wave per job spends its time on memory round trips and on
// per-job counter atomics.  Here each lane runs one such job (at most
// kSmallCut nodes) -- the three stable sorts, minCost and the refinement --
// sequentially, with the same FP64 operations in the same order as the wave
// form (min/max unions are exact in any order), so the tree is the same; the
// wave then allocates the outputs of its 64 jobs with one atomic per counter.

__device__ inline double centre(const Flat& f, int32_t node, int axis) {
  const float* b = f.box + 6 * size_t(node);
  return (double(b[axis]) + double(b[3 + axis])) * (1.0 / 2.0);
}

// Stable sort by box centre on `axis` of the members in order `in` (indices
// into ids): insertion sort on (centre, input position).
template <int N = kSmallCut>
__device__ inline void small_sort(const Flat& f, const int32_t* ids, const int* in, int* o, int m, int axis) {
  double key[N];
  for (int i = 0; i < N; ++i) {
    if (i >= m) break;
    o[i] = in ? in[i] : i;
    key[i] = centre(f, ids[o[i]], axis);
    for (int j = i; j > 0 && key[j - 1] > key[j]; --j) {
      const double tk = key[j - 1];
      key[j - 1] = key[j];
      key[j] = tk;
      const int to = o[j - 1];
      o[j - 1] = o[j];
      o[j] = to;
    }
  }
}

__device__ inline void box_merge(float* lo, float* hi, const Flat& f, int32_t node) {
  const float* b = f.box + 6 * size_t(node);
  for (int a = 0; a < 3; ++a) {
    lo[a] = fminf(lo[a], b[a]);
    hi[a] = fmaxf(hi[a], b[3 + a]);
  }
}

__device__ inline double sa_lohi(const float* lo, const float* hi) {
  const float bb[6] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
  return sa_box(bb);
}

// minCost (phr.go:414-459) of order o, as min_cost computes it.
template <int N = kSmallCut>
__device__ inline Split small_cost(const Flat& f, const int32_t* ids, const int* o, int m, const float* cut_box) {
#pragma clang fp contract(off)
  double rc[N];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t cnt = 0;
  for (int r = m - 1; r >= 1; --r) {
    const int32_t node = ids[o[r]];
    box_merge(lo, hi, f, node);
    cnt += uint32_t(f.size[node]);
    rc[r] = sa_lohi(lo, hi) * double(cnt);
  }
  const uint32_t total = cnt + uint32_t(f.size[ids[o[0]]]);
  float plo[3] = {INFINITY, INFINITY, INFINITY}, phi[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t pcnt = 0;
  double best = INFINITY;
  int32_t best_at = m;
  for (int r = 1; r < m; ++r) {
    const int32_t node = ids[o[r - 1]];
    box_merge(plo, phi, f, node);
    pcnt += uint32_t(f.size[node]);
    const double c = rc[r] + sa_lohi(plo, phi) * double(pcnt);
    if (c < best) {
      best = c;
      best_at = r;
    }
  }
  const double nosplit = sa_box(cut_box) * double(total);
  if (best < nosplit) return Split{best, best_at};
  return Split{nosplit, 0};
}

__device__ inline bool keep_node(const Flat& f, int32_t node, double th) {
  return f.child[2 * node] < 0 || sa_box(f.box + 6 * size_t(node)) < th;
}

__device__ inline uint32_t wave_max(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, uint32_t(__shfl_xor(v, o, 64)));
  return v;
}

// Wave-aggregated allocation: every lane receives base + its exclusive prefix.
__device__ inline uint32_t wave_alloc(uint32_t* counter, uint32_t n) {
  uint32_t total;
  const uint32_t ex = wave_excl(n, &total);
  uint32_t base = 0;
  if ((threadIdx.x & 63) == 0 && total) base = atomicAdd(counter, total);
  return __shfl(base, 0, 64) + ex;
}

// One job per lane; every lane of the wave calls it (the allocations are
// wave-wide).
__device__ void small_job(int32_t job, Flat f, const double* __restrict__ th, const Task* __restrict__ tasks,
                          int32_t n_jobs, const int32_t* __restrict__ pool, int32_t* __restrict__ pool_next,
                          Lists next, OutNode* __restrict__ out, int32_t* __restrict__ list, Ctl* ctl, Caps caps) {
  bool valid = job < n_jobs;
  Task t{};
  int m = 0;
  if (valid) {
    t = tasks[job];
    m = t.cnt;
    if (m < 2 || m > kSmallCut || t.depth >= kMaxDepth) {
      atomicOr(&ctl->err, 2u);
      valid = false;
    }
  }
  int32_t ids[kSmallCut];
  int oX[kSmallCut], oY[kSmallCut], oZ[kSmallCut];
  const int* C = oX;
  Split s{0.0, 0};
  // what the job emits: output nodes, leaf-list entries, and per half a leaf
  // (k == 1) or a job of k pool entries
  uint32_t n_out = 0, n_list = 0, k_half[2] = {0, 0};
  uint32_t pl = 0, pr = 0, pc = 0;
  float L[6], R[6];
  double thd = 0.0;
  if (valid) {
    for (int k = 0; k < kSmallCut; ++k)
      if (k < m) ids[k] = pool[t.off + k];
    // SweepSAH (phr.go:368-405)
    small_sort(f, ids, nullptr, oX, m, 0);
    small_sort(f, ids, oX, oY, m, 1);
    const Split x = small_cost(f, ids, oX, m, t.box);
    const Split y = small_cost(f, ids, oY, m, t.box);
    small_sort(f, ids, x.cost < y.cost ? oY : oX, oZ, m, 2);
    const Split z = small_cost(f, ids, oZ, m, t.box);
    if (x.cost < y.cost) {
      if (x.cost < z.cost) { C = oX; s = x; } else { C = oZ; s = z; }
    } else {
      if (y.cost < z.cost) { C = oY; s = y; } else { C = oZ; s = z; }
    }
    if (s.at > 0) {
      for (int a = 0; a < 3; ++a) {
        L[a] = R[a] = INFINITY;
        L[3 + a] = R[3 + a] = -INFINITY;
      }
      thd = th[t.depth];  // refined() uses the split cut's depth
      for (int r = 0; r < m; ++r) {
        const int32_t node = ids[C[r]];
        const uint32_t p = uint32_t(f.prim[2 * node + 1] - f.prim[2 * node]);
        const uint32_t k = keep_node(f, node, thd) ? 1u : 2u;
        if (r < s.at) {
          box_merge(L, L + 3, f, node);
          pl += p;
          k_half[0] += k;
        } else {
          box_merge(R, R + 3, f, node);
          pr += p;
          k_half[1] += k;
        }
      }
      n_out = 2;
      for (int h = 0; h < 2; ++h) n_list += k_half[h] == 1 ? 1u : 0u;
    } else {
      for (int r = 0; r < m; ++r) {
        const int32_t node = ids[C[r]];
        pc += uint32_t(f.prim[2 * node + 1] - f.prim[2 * node]);
      }
      n_out = 1;
      n_list = uint32_t(m);
    }
  }
  // one allocation per counter per wave
  uint32_t n_pool = 0, n_job[kLists] = {0, 0, 0, 0}, max_lg = 0, max_lvl = 0;
  for (int h = 0; h < 2; ++h) {
    const uint32_t k = k_half[h];
    if (k >= 2) {
      n_pool += k;
      const int dst = route(k);
      ++n_job[dst];
      if (dst == kLarge) max_lg = max(max_lg, k);
    }
  }
  if (valid && (s.at == 0 || k_half[0] == 1 || k_half[1] == 1)) max_lvl = uint32_t(t.level + 1);
  const int32_t o = int32_t(wave_alloc(&ctl->out, n_out));
  const int32_t ls = int32_t(wave_alloc(&ctl->list, n_list));
  const int32_t po = int32_t(wave_alloc(&ctl->pool, n_pool));
  int32_t slot[kLists];
  for (int r = 0; r < kLists; ++r) slot[r] = int32_t(wave_alloc(&ctl->jobs[r][0], n_job[r]));
  max_lg = wave_max(max_lg);
  max_lvl = wave_max(max_lvl);
  if ((threadIdx.x & 63) == 0) {
    if (max_lg) raise_max(&ctl->max_m, max_lg);
    if (max_lvl) raise_max(&ctl->max_level, max_lvl);
  }
  if (!valid) return;
  bool over = int64_t(o) + n_out > caps.out || int64_t(ls) + n_list > caps.list || int64_t(po) + n_pool > caps.pool;
  for (int r = 0; r < kLists; ++r) over = over || int64_t(slot[r]) + n_job[r] > next.cap[r];
  if (over) {
    atomicOr(&ctl->err, 1u);
    return;
  }
  if (s.at == 0) {
    // no split beats the whole cut: one leaf of every triangle, in order C
    for (int r = 0; r < m; ++r) list[ls + r] = ids[C[r]];
    write_out(out, o, 1, 0, -1, -1, ls, m, t.prim_off, int32_t(pc), t.level + 1, t.box);
    write_out(out, t.out, 0, 1, o, -1, -1, 0, t.prim_off, int32_t(pc), t.level, t.box);
    return;
  }
  write_out(out, t.out, 0, 2, o, o + 1, -1, 0, t.prim_off, int32_t(pl + pr), t.level, t.box);
  int32_t lnext = ls, pnext = po;
  for (int h = 0; h < 2; ++h) {
    const int a = h == 0 ? 0 : s.at, b = h == 0 ? s.at : m;
    const float* hb = h == 0 ? L : R;
    const int32_t poff_prim = h == 0 ? t.prim_off : t.prim_off + int32_t(pl);
    const int32_t pcnt = int32_t(h == 0 ? pl : pr);
    const uint32_t k = k_half[h];
    if (k == 1) {  // one kept node: a leaf of all its triangles
      list[lnext] = ids[C[a]];
      write_out(out, o + h, 1, 0, -1, -1, lnext, 1, poff_prim, pcnt, t.level + 1, hb);
      ++lnext;
      continue;
    }
    int32_t p = pnext;
    for (int r = a; r < b; ++r) {
      const int32_t node = ids[C[r]];
      if (keep_node(f, node, thd)) {
        pool_next[p++] = node;
      } else {
        pool_next[p++] = f.child[2 * node];
        pool_next[p++] = f.child[2 * node + 1];
      }
    }
    Task n;
    n.off = pnext;
    n.cnt = int32_t(k);
    for (int q = 0; q < 6; ++q) n.box[q] = hb[q];
    n.depth = t.depth + 1;
    n.out = o + h;
    n.prim_off = poff_prim;
    n.level = t.level + 1;
    const int dst = route(k);
    next.p[dst][slot[dst]++] = n;
    pnext += int32_t(k);
  }
}

// ---- register-resident jobs: G lanes per job -------------------------------------
// Jobs with cuts of kSmallCut+1 .. G nodes: each lane of a group of G holds one
// member (node id, box, subtree size, triangle count), loaded once.  The sorts
// are bitonic networks over the group's lanes on (centre, position) pairs
// followed by one gather of the records, the minCost sweeps are lane scans,
// and outputs are allocated with one atomic per counter per wave -- the same
// FP64 operations as the staged wave form, with no memory round trips between
// the steps.

struct Mem {
  int32_t node;
  float lo[3], hi[3];
  uint32_t size, prims;
};

__device__ inline Mem mem_shfl(const Mem& a, int src) {
  Mem r;
  r.node = __shfl(a.node, src, 64);
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = __shfl(a.lo[k], src, 64);
    r.hi[k] = __shfl(a.hi[k], src, 64);
  }
  r.size = __shfl(a.size, src, 64);
  r.prims = __shfl(a.prims, src, 64);
  return r;
}

__device__ inline WBox wb_shfl_down(const WBox& b, int o) {
  WBox r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = __shfl_down(b.lo[k], o, 64);
    r.hi[k] = __shfl_down(b.hi[k], o, 64);
  }
  r.cnt = __shfl_down(b.cnt, o, 64);
  return r;
}

// sort.SliceStable of the group's members by box centre on `axis` (lanes at
// or beyond m carry +inf and stay behind the members).
template <int G>
__device__ inline Mem group_sort(const Mem& a, int g, int gbase, int m, int axis) {
  double key = g < m ? (double(a.lo[axis]) + double(a.hi[axis])) * (1.0 / 2.0) : INFINITY;
  int32_t pos = g;
  for (int k = 2; k <= G; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      const double ok = __shfl_xor(key, j, 64);
      const int32_t op = __shfl_xor(pos, j, 64);
      const bool other_less = ok < key || (ok == key && op < pos);
      const bool keep_min = ((g & j) == 0) == ((g & k) == 0);
      if (keep_min == other_less) {
        key = ok;
        pos = op;
      }
    }
  }
  return mem_shfl(a, gbase + pos);
}

template <int G>
__device__ inline WBox group_reduce(WBox v) {
  for (int o = G / 2; o > 0; o >>= 1) wb_merge(v, wb_shfl_xor(v, o));
  return v;
}

template <int G>
__device__ inline uint32_t group_sum(uint32_t v) {
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// minCost (phr.go:414-459) of the group's members in lane order, as min_cost
// computes it.
template <int G>
__device__ inline Split group_cost(const Mem& a, int g, int gbase, int m, const float* cut_box) {
#pragma clang fp contract(off)
  WBox v;
  wb_empty(v);
  if (g < m) {
    for (int k = 0; k < 3; ++k) {
      v.lo[k] = a.lo[k];
      v.hi[k] = a.hi[k];
    }
    v.cnt = a.size;
  }
  WBox R = v;  // union of [g, m)
  for (int o = 1; o < G; o <<= 1) {
    const WBox u = wb_shfl_down(R, o);
    if (g + o < G) wb_merge(R, u);
  }
  WBox I = v;  // union of [0, g]
  for (int o = 1; o < G; o <<= 1) {
    const WBox u = wb_shfl_up(I, o);
    if (g >= o) wb_merge(I, u);
  }
  WBox P = wb_shfl_up(I, 1);  // union of [0, g)
  if (g == 0) wb_empty(P);
  double best = INFINITY;
  int32_t best_at = m;
  if (g >= 1 && g < m) {
    const double rc = sa_wb(R) * double(R.cnt);
    best = rc + sa_wb(P) * double(P.cnt);
    best_at = g;
  }
  for (int o = G / 2; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64);
    const int32_t oa = __shfl_xor(best_at, o, 64);
    if (ob < best || (ob == best && oa < best_at)) {
      best = ob;
      best_at = oa;
    }
  }
  const uint32_t total = __shfl(R.cnt, gbase, 64);
  const double nosplit = sa_box(cut_box) * double(total);
  if (best < nosplit) return Split{best, best_at};
  return Split{nosplit, 0};
}

// One job per G lanes; every lane of the wave calls it.
template <int G>
__device__ void group_job(int32_t job, Flat f, const double* __restrict__ th, const Task* __restrict__ tasks,
                          int32_t n_jobs, const int32_t* __restrict__ pool, int32_t* __restrict__ pool_next,
                          Lists next, OutNode* __restrict__ out, int32_t* __restrict__ list, Ctl* ctl, Caps caps) {
  static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0, "groups are powers of two within a wave");
  const int lane = threadIdx.x & 63;
  const int g = lane & (G - 1);
  const int gbase = lane - g;
  const bool lead = g == 0;
  // every lane of a group reads the same job, so control flow stays uniform
  // per group; all cross-lane steps below run with the whole wave active
  bool valid = job < n_jobs;
  Task t{};
  int m = 0;
  if (valid) {
    t = tasks[job];
    m = t.cnt;
    if (m < 2 || m > G || t.depth >= kMaxDepth) {
      if (lead) atomicOr(&ctl->err, 2u);
      valid = false;
      m = 0;
    }
  }
  Mem a;
  a.node = -1;
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = INFINITY;
    a.hi[k] = -INFINITY;
  }
  a.size = 0;
  a.prims = 0;
  if (g < m) {
    a.node = pool[t.off + g];
    const float* b = f.box + 6 * size_t(a.node);
    for (int k = 0; k < 3; ++k) {
      a.lo[k] = b[k];
      a.hi[k] = b[3 + k];
    }
    a.size = uint32_t(f.size[a.node]);
    a.prims = uint32_t(f.prim[2 * a.node + 1] - f.prim[2 * a.node]);
  }

  // SweepSAH (phr.go:368-405)
  const Mem X = group_sort<G>(a, g, gbase, m, 0);
  const Mem Y = group_sort<G>(X, g, gbase, m, 1);
  const Split sx = group_cost<G>(X, g, gbase, m, t.box);
  const Split sy = group_cost<G>(Y, g, gbase, m, t.box);
  const bool x_first = sx.cost < sy.cost;
  const Mem Z = group_sort<G>(x_first ? Y : X, g, gbase, m, 2);
  const Split sz = group_cost<G>(Z, g, gbase, m, t.box);
  Mem c;
  Split s;
  if (x_first) {
    if (sx.cost < sz.cost) { c = X; s = sx; } else { c = Z; s = sz; }
  } else {
    if (sy.cost < sz.cost) { c = Y; s = sy; } else { c = Z; s = sz; }
  }

  // refined() (phr.go:313-338) with the split cut's depth
  const bool member = g < m;
  const bool left = member && g < s.at;
  const bool right = member && s.at > 0 && g >= s.at;
  uint32_t k = 0;
  bool keep = true;
  if (s.at > 0 && member) {
    const double thd = th[t.depth];
    const float bb[6] = {c.lo[0], c.lo[1], c.lo[2], c.hi[0], c.hi[1], c.hi[2]};
    keep = f.child[2 * c.node] < 0 || sa_box(bb) < thd;
    k = keep ? 1u : 2u;
  }
  const uint32_t kL = group_sum<G>(left ? k : 0u);
  const uint32_t kR = group_sum<G>(right ? k : 0u);
  const uint32_t pl = group_sum<G>(left ? c.prims : 0u);
  const uint32_t pr = group_sum<G>(right ? c.prims : 0u);
  const uint32_t pall = group_sum<G>(member ? c.prims : 0u);
  WBox hv;
  wb_empty(hv);
  WBox hw = hv;
  if (left || right) {
    WBox& dst = left ? hv : hw;
    for (int q = 0; q < 3; ++q) {
      dst.lo[q] = c.lo[q];
      dst.hi[q] = c.hi[q];
    }
  }
  const WBox Lb = group_reduce<G>(hv);
  const WBox Rb = group_reduce<G>(hw);
  uint32_t kx = k;  // exclusive scan of k over the group
  for (int o = 1; o < G; o <<= 1) {
    const uint32_t y = __shfl_up(kx, o, 64);
    if (g >= o) kx += y;
  }
  kx -= k;
  const int32_t first_left = __shfl(c.node, gbase, 64);
  const int32_t first_right = __shfl(c.node, gbase + (s.at > 0 ? s.at : 0), 64);

  // what the job emits; the group leader's needs are allocated per wave
  uint32_t n_out = 0, n_list = 0, n_pool = 0, n_job[kLists] = {0, 0, 0, 0}, max_lg = 0, max_lvl = 0;
  if (valid) {
    if (s.at > 0) {
      n_out = 2;
      for (int h = 0; h < 2; ++h) {
        const uint32_t kh = h ? kR : kL;
        if (kh == 1) {
          ++n_list;
          max_lvl = uint32_t(t.level + 1);
        } else {
          n_pool += kh;
          const int dst = route(kh);
          ++n_job[dst];
          if (dst == kLarge) max_lg = max(max_lg, kh);
        }
      }
    } else {
      n_out = 1;
      n_list = uint32_t(m);
      max_lvl = uint32_t(t.level + 1);
    }
  }
  const int32_t o = __shfl(int32_t(wave_alloc(&ctl->out, lead ? n_out : 0u)), gbase, 64);
  const int32_t ls = __shfl(int32_t(wave_alloc(&ctl->list, lead ? n_list : 0u)), gbase, 64);
  const int32_t po = __shfl(int32_t(wave_alloc(&ctl->pool, lead ? n_pool : 0u)), gbase, 64);
  int32_t slot[kLists];
  for (int r = 0; r < kLists; ++r)
    slot[r] = __shfl(int32_t(wave_alloc(&ctl->jobs[r][0], lead ? n_job[r] : 0u)), gbase, 64);
  max_lg = wave_max(lead ? max_lg : 0u);
  max_lvl = wave_max(lead ? max_lvl : 0u);
  if (lane == 0) {
    if (max_lg) raise_max(&ctl->max_m, max_lg);
    if (max_lvl) raise_max(&ctl->max_level, max_lvl);
  }
  if (!valid) return;
  bool over = int64_t(o) + n_out > caps.out || int64_t(ls) + n_list > caps.list || int64_t(po) + n_pool > caps.pool;
  for (int r = 0; r < kLists; ++r) over = over || int64_t(slot[r]) + n_job[r] > next.cap[r];
  if (over) {
    if (lead) atomicOr(&ctl->err, 1u);
    return;
  }
  if (s.at == 0) {
    // no split beats the whole cut: one leaf of every triangle, in sorted order
    if (member) list[ls + g] = c.node;
    if (lead) {
      write_out(out, o, 1, 0, -1, -1, ls, m, t.prim_off, int32_t(pall), t.level + 1, t.box);
      write_out(out, t.out, 0, 1, o, -1, -1, 0, t.prim_off, int32_t(pall), t.level, t.box);
    }
    return;
  }
  if (lead) write_out(out, t.out, 0, 2, o, o + 1, -1, 0, t.prim_off, int32_t(pl + pr), t.level, t.box);
  int32_t lnext = ls, pnext = po;
  for (int h = 0; h < 2; ++h) {
    const uint32_t kh = h ? kR : kL;
    const WBox& hb = h ? Rb : Lb;
    const float box[6] = {hb.lo[0], hb.lo[1], hb.lo[2], hb.hi[0], hb.hi[1], hb.hi[2]};
    const int32_t poff_prim = h ? t.prim_off + int32_t(pl) : t.prim_off;
    const int32_t pcnt = int32_t(h ? pr : pl);
    if (kh == 1) {  // one kept node: a leaf of all its triangles
      if (lead) {
        list[lnext] = h ? first_right : first_left;
        write_out(out, o + h, 1, 0, -1, -1, lnext, 1, poff_prim, pcnt, t.level + 1, box);
      }
      ++lnext;
      continue;
    }
    if (h ? right : left) {
      const int32_t p = pnext + int32_t(kx - (h ? kL : 0u));
      if (keep) {
        pool_next[p] = c.node;
      } else {
        pool_next[p] = f.child[2 * c.node];
        pool_next[p + 1] = f.child[2 * c.node + 1];
      }
    }
    if (lead) {
      Task n;
      n.off = pnext;
      n.cnt = int32_t(kh);
      for (int q = 0; q < 6; ++q) n.box[q] = box[q];
      n.depth = t.depth + 1;
      n.out = o + h;
      n.prim_off = poff_prim;
      n.level = t.level + 1;
      const int dst = route(kh);
      next.p[dst][slot[dst]++] = n;
    }
    pnext += int32_t(kh);
  }
}

// Device-driven levels: at the start of a level, the counts its jobs appended
// become the current ones (the kernels of the level read those) and the
// counters restart for the next level; BuildWithCost's counter (phr.go:221)
// gains the level's pool.  Also picks the large jobs' workgroup size and LDS
// sort size: 512 threads for a few jobs of big cuts and for cuts that need
// more than kSmallLds of bitonic sort, 64 for thousands of jobs, else 256.
// The words are read and written with device-scope atomics: this runs at the
// end of a level's last launch, whose other workgroups updated them.
__device__ inline uint32_t ctl_ld(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void ctl_st(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ void advance_level(Ctl* ctl) {
  uint32_t total = 0;
  for (int r = 0; r < kLists; ++r) {
    const uint32_t c = ctl_ld(&ctl->jobs[r][0]);
    ctl_st(&ctl->cur_jobs[r], c);
    total += c;
    ctl_st(&ctl->jobs[r][0], 0u);
  }
  const uint32_t mm = ctl_ld(&ctl->max_m);
  const uint32_t nb = ctl_ld(&ctl->cur_jobs[kLarge]);
  uint32_t lds_p = 0;
  if (mm > 64u && mm <= uint32_t(kSortLds)) {
    lds_p = 128;
    while (lds_p < mm) lds_p <<= 1;
  }
  uint32_t bsz = 64;
  if (mm > 64u && nb < 2048u) bsz = ((nb <= 128u && mm > 256u) || lds_p > uint32_t(kSmallLds)) ? 512u : 256u;
  ctl_st(&ctl->cur_max_m, mm);
  ctl_st(&ctl->cur_lds, lds_p);
  ctl_st(&ctl->sel_b, bsz);
  if (total > 0) {
    const uint32_t pool = ctl_ld(&ctl->pool);
    __hip_atomic_fetch_add(&ctl->work, (unsigned long long)pool, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ctl_st(&ctl->levels, ctl_ld(&ctl->levels) + 1u);
  }
  ctl_st(&ctl->pool, 0u);
  ctl_st(&ctl->max_m, 0u);
  ctl_st(&ctl->groups, 0u);
}

__global__ void advance_level_kernel(Ctl* ctl) { advance_level(ctl); }

// ---- n-ary groups whose cuts are small: one lane per group ----------------------------

// A group whose cut to split has at most kNaryLaneCut members (it is the group's
// largest cut, so every cut of the group is that small) takes one lane: the
// split is the lane job's (same sorts, minCost and refinement, the same FP64
// operations), then the group's bookkeeping as nary_kernel does it -- the
// split cut replaced by its first half, the second appended, the waiting cuts'
// members carried into the next pool, then either the group moves on (fewer
// than bf cuts and one of two or more members) or its branch is written.
// Every lane of the wave calls it: the allocations are wave-wide.
__device__ void nary_lane_job(int32_t gi, int32_t n_groups, Flat f, const double* __restrict__ th,
                              const Group* __restrict__ groups, Group* __restrict__ groups_next,
                              const int32_t* __restrict__ pool, int32_t* __restrict__ pool_next,
                              OutNode* __restrict__ out, int32_t* __restrict__ list, Ctl* ctl, NCaps caps, int bf) {
  bool valid = gi < n_groups;
  const Group* gp = groups + (caps.groups - 1 - (valid ? gi : 0));  // the small list runs from the back
  int32_t ncuts = 0, si = 0;
  CutRec c{};
  int m = 0;
  if (valid) {
    ncuts = gp->ncuts;
    si = gp->split;
    if (si >= 0 && si < ncuts && ncuts <= kMaxBranching) c = gp->cuts[si];
    m = c.cnt;
    if (c.depth >= kMaxDepth || si < 0 || si >= ncuts || ncuts > kMaxBranching || m < 2 || m > kNaryLaneCut ||
        c.leaf_lc > 0) {
      atomicOr(&ctl->err, 2u);
      valid = false;
      m = 0;
    }
  }
  int32_t ids[kNaryLaneCut];
  int oX[kNaryLaneCut], oY[kNaryLaneCut], oZ[kNaryLaneCut];
  const int* C = oX;
  Split sp{0.0, 0};
  uint32_t k0 = 0, k1 = 0, pl = 0, pr = 0;
  float L[6], R[6];
  double thd = 0.0;
  if (valid) {
    for (int k = 0; k < kNaryLaneCut; ++k)
      if (k < m) ids[k] = pool[c.off + k];
    // SweepSAH (phr.go:368-405)
    small_sort<kNaryLaneCut>(f, ids, nullptr, oX, m, 0);
    small_sort<kNaryLaneCut>(f, ids, oX, oY, m, 1);
    const Split x = small_cost<kNaryLaneCut>(f, ids, oX, m, c.box);
    const Split y = small_cost<kNaryLaneCut>(f, ids, oY, m, c.box);
    small_sort<kNaryLaneCut>(f, ids, x.cost < y.cost ? oY : oX, oZ, m, 2);
    const Split z = small_cost<kNaryLaneCut>(f, ids, oZ, m, c.box);
    if (x.cost < y.cost) {
      if (x.cost < z.cost) { C = oX; sp = x; } else { C = oZ; sp = z; }
    } else {
      if (y.cost < z.cost) { C = oY; sp = y; } else { C = oZ; sp = z; }
    }
    if (sp.at > 0) {
      for (int a = 0; a < 3; ++a) {
        L[a] = R[a] = INFINITY;
        L[3 + a] = R[3 + a] = -INFINITY;
      }
      thd = th[c.depth];  // refined() uses the split cut's depth
      for (int r = 0; r < m; ++r) {
        const int32_t node = ids[C[r]];
        const uint32_t pc = uint32_t(f.prim[2 * node + 1] - f.prim[2 * node]);
        const uint32_t k = keep_node(f, node, thd) ? 1u : 2u;
        if (r < sp.at) {
          box_merge(L, L + 3, f, node);
          pl += pc;
          k0 += k;
        } else {
          box_merge(R, R + 3, f, node);
          pr += pc;
          k1 += k;
        }
      }
    }
  }
  const bool two = valid && sp.at > 0;
  // the group's cuts after the split: member counts (a leaf counts 1) and leaf flags
  const int n_new = valid ? ncuts + (two ? 1 : 0) : 0;
  uint32_t cnt[kMaxBranching];
  bool leaf[kMaxBranching];
  uint32_t carried = 0;
#pragma unroll
  for (int j = 0; j < kMaxBranching; ++j) {
    cnt[j] = 0;
    leaf[j] = false;
    if (j >= n_new) continue;
    if (j == si) {
      leaf[j] = !two || k0 == 1;
      cnt[j] = leaf[j] ? 1u : k0;
    } else if (j == ncuts) {  // the appended second half
      leaf[j] = k1 == 1;
      cnt[j] = leaf[j] ? 1u : k1;
    } else {
      const CutRec cj = gp->cuts[j];
      leaf[j] = cj.leaf_lc > 0;
      cnt[j] = uint32_t(cj.cnt);
      if (!leaf[j]) carried += cnt[j];
    }
  }
  uint32_t mx = 0;
  int mi = 0;
#pragma unroll
  for (int j = 0; j < kMaxBranching; ++j)
    if (j < n_new && cnt[j] > mx) {
      mx = cnt[j];
      mi = j;
    }
  const bool cont = valid && n_new < bf && mx > 1;
  const bool branch = valid && !cont;
  // what the lane emits
  const uint32_t n_list = !valid ? 0u : !two ? uint32_t(m) : (k0 == 1 ? 1u : 0u) + (k1 == 1 ? 1u : 0u);
  const uint32_t n_pool = !valid ? 0u : (two && k0 > 1 ? k0 : 0u) + (two && k1 > 1 ? k1 : 0u) + carried;
  const uint32_t n_out = branch ? uint32_t(n_new) : 0u;
  uint32_t g_big = 0, g_small = 0, max_big = 0, max_lvl = 0;
  if (cont) {
    if (mx <= uint32_t(kNaryLaneCut)) g_small = 1;
    else {
      g_big = 1;
      max_big = mx;
    }
  }
  if (branch) {
#pragma unroll
    for (int j = 0; j < kMaxBranching; ++j) {
      if (j >= n_new) continue;
      if (leaf[j]) {
        max_lvl = 1;
      } else if (cnt[j] <= uint32_t(kNaryLaneCut)) {
        ++g_small;
      } else {
        ++g_big;
        max_big = max(max_big, cnt[j]);
      }
    }
  }
  const int32_t ls = int32_t(wave_alloc(&ctl->list, n_list));
  const int32_t po = int32_t(wave_alloc(&ctl->pool, n_pool));
  const int32_t o = int32_t(wave_alloc(&ctl->out, n_out));
  const int32_t gb = int32_t(wave_alloc(&ctl->groups, g_big));
  const int32_t gs = int32_t(wave_alloc(&ctl->groups_small, g_small));
  // buildSubTreeCost's counter (phr.go:221) and the level bounds
  uint32_t wsum;
  wave_excl(uint32_t(m), &wsum);
  const uint32_t lvl = max_lvl ? uint32_t(gp->level + 1) : 0u;
  const uint32_t w_big = wave_max(max_big), w_lvl = wave_max(lvl);
  if ((threadIdx.x & 63) == 0) {
    if (wsum) atomicAdd(&ctl->work, (unsigned long long)wsum);
    if (w_big) raise_max(&ctl->max_m, w_big);
    if (w_lvl) raise_max(&ctl->max_level, w_lvl);
  }
  if (!valid) return;
  if (int64_t(ls) + n_list > caps.list || int64_t(po) + n_pool > caps.pool || int64_t(o) + n_out > caps.out ||
      int64_t(gb) + g_big > caps.groups || int64_t(gs) + g_small > caps.groups) {
    atomicOr(&ctl->err, 1u);
    return;
  }
  // the two halves (or the leaf of the whole cut) and their pool entries
  CutRec first = c, second = c;
  int32_t lnext = ls, pnext = po;
  if (!two) {
    for (int r = 0; r < m; ++r) list[ls + r] = ids[C[r]];
    first.off = 0;
    first.cnt = 1;
    first.leaf_ls = ls;
    first.leaf_lc = m;
  } else {
    for (int h = 0; h < 2; ++h) {
      CutRec& rec = h ? second : first;
      const int a = h ? sp.at : 0, b = h ? m : sp.at;
      const uint32_t k = h ? k1 : k0;
      rec.depth = c.depth + 1;
      rec.prims = int32_t(h ? pr : pl);
      for (int q = 0; q < 6; ++q) rec.box[q] = h ? R[q] : L[q];
      if (k == 1) {
        list[lnext] = ids[C[a]];
        rec.off = 0;
        rec.cnt = 1;
        rec.leaf_ls = lnext;
        rec.leaf_lc = 1;
        ++lnext;
        continue;
      }
      rec.off = pnext;
      rec.cnt = int32_t(k);
      rec.leaf_ls = 0;
      rec.leaf_lc = 0;
      for (int r = a; r < b; ++r) {
        const int32_t node = ids[C[r]];
        if (keep_node(f, node, thd)) {
          pool_next[pnext++] = node;
        } else {
          pool_next[pnext++] = f.child[2 * node];
          pool_next[pnext++] = f.child[2 * node + 1];
        }
      }
    }
  }
  const int32_t g_out = gp->out, g_level = gp->level, g_prim = gp->prim_off;
  float g_box[6];
  for (int q = 0; q < 6; ++q) g_box[q] = gp->box[q];
  int32_t big_next = gb, small_next = gs;
  Group* dst = nullptr;
  if (cont) {
    dst = mx <= uint32_t(kNaryLaneCut) ? groups_next + (caps.groups - 1 - small_next++) : groups_next + big_next++;
    dst->out = g_out;
    dst->level = g_level;
    dst->prim_off = g_prim;
    dst->ncuts = n_new;
    dst->split = mi;
    for (int q = 0; q < 3; ++q) dst->pad[q] = 0;
    for (int q = 0; q < 6; ++q) dst->box[q] = g_box[q];
    dst->pad2[0] = dst->pad2[1] = 0;
  }
  int32_t p = g_prim;
  for (int j = 0; j < n_new; ++j) {
    CutRec cj;
    if (j == si) {
      cj = first;
    } else if (j == ncuts) {
      cj = second;
    } else {
      cj = gp->cuts[j];
      if (cj.leaf_lc == 0) {  // a waiting cut: its members move to the next pool
        for (int r = 0; r < cj.cnt; ++r) pool_next[pnext + r] = pool[cj.off + r];
        cj.off = pnext;
        pnext += cj.cnt;
      }
    }
    if (cont) {
      dst->cuts[j] = cj;
      continue;
    }
    // the branch's child j
    if (cj.leaf_lc > 0) {
      write_out(out, o + j, 1, 0, -1, -1, cj.leaf_ls, cj.leaf_lc, p, cj.prims, g_level + 1, cj.box);
    } else {
      Group* ng = uint32_t(cj.cnt) <= uint32_t(kNaryLaneCut) ? groups_next + (caps.groups - 1 - small_next++)
                                                          : groups_next + big_next++;
      ng->out = o + j;
      ng->level = g_level + 1;
      ng->prim_off = p;
      ng->ncuts = 1;
      ng->split = 0;
      for (int q = 0; q < 3; ++q) ng->pad[q] = 0;
      for (int q = 0; q < 6; ++q) ng->box[q] = cj.box[q];
      ng->pad2[0] = ng->pad2[1] = 0;
      ng->cuts[0] = cj;
    }
    p += cj.prims;
  }
  if (branch)
    write_out(out, g_out, 0, n_new, o, n_new > 1 ? o + 1 : -1, -1, 0, g_prim, p - g_prim, g_level, g_box);
}

__global__ __launch_bounds__(256) void nary_lane_kernel(Flat f, const double* __restrict__ th,
                                                        const Group* __restrict__ groups, int32_t n_groups,
                                                        Group* __restrict__ groups_next,
                                                        const int32_t* __restrict__ pool,
                                                        int32_t* __restrict__ pool_next, OutNode* __restrict__ out,
                                                        int32_t* __restrict__ list, Ctl* ctl, NCaps caps, int bf) {
  nary_lane_job(int32_t(blockIdx.x * 256 + threadIdx.x), n_groups, f, th, groups, groups_next, pool, pool_next, out,
                list, ctl, caps, bf);
}

// Small and group jobs of the current level, grid-stride over its count.
// The small jobs run last in a level: the workgroup that finishes last
// (arrival counter ctl->done) advances the counters for the next level.
__global__ __launch_bounds__(256) void small_task_kernel(Flat f, const double* __restrict__ th,
                                                         const Task* __restrict__ tasks,
                                                         const int32_t* __restrict__ pool,
                                                         int32_t* __restrict__ pool_next, Lists next,
                                                         OutNode* __restrict__ out, int32_t* __restrict__ list,
                                                         Ctl* ctl, Caps caps) {
  __shared__ uint32_t s_last;
  const int32_t n_jobs = int32_t(ctl->cur_jobs[0]);
  for (int64_t base = blockIdx.x * int64_t(256); base < n_jobs; base += int64_t(gridDim.x) * 256)
    small_job(int32_t(base) + int32_t(threadIdx.x), f, th, tasks, n_jobs, pool, pool_next, next, out, list, ctl,
              caps);
  // this workgroup's updates of the control block are device-scope atomics;
  // waiting for their completion orders them before its arrival, and the
  // last arrival reads the block with device-scope loads
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0)
    s_last = atomicAdd(&ctl->done, 1u) == gridDim.x - 1 ? 1u : 0u;
  __syncthreads();
  if (s_last && threadIdx.x == 0) {
    ctl_st(&ctl->done, 0u);
    advance_level(ctl);
  }
}

template <int G>
__global__ __launch_bounds__(256) void group_task_kernel(Flat f, const double* __restrict__ th,
                                                         const Task* __restrict__ tasks,
                                                         const int32_t* __restrict__ pool,
                                                         int32_t* __restrict__ pool_next, Lists next,
                                                         OutNode* __restrict__ out, int32_t* __restrict__ list,
                                                         Ctl* ctl, Caps caps) {
  const int32_t n_jobs = int32_t(ctl->cur_jobs[G == kWaveCut ? 2 : 1]);
  constexpr int kPerBlock = 256 / G;
  for (int64_t base = blockIdx.x * int64_t(kPerBlock); base < n_jobs; base += int64_t(gridDim.x) * kPerBlock)
    group_job<G>(int32_t(base) + int32_t(threadIdx.x / G), f, th, tasks, n_jobs, pool, pool_next, next, out, list,
                 ctl, caps);
}

// The first device-driven level: the initial cut's job (on the large list)
// and its pool.
// d_err (a rebuild recorded in a graph): an initial cut of one node, which an
// eager build answers by attaching the LBVH as is, is out of the recorded
// bounds; no job starts and the error word says so.
__global__ void seed_levels_kernel(Ctl* ctl, uint32_t* d_err) {
  const bool run = !d_err || ctl->cut >= 2u;
  if (!run) {
    atomicOr(d_err, TRACE_ERR_BOUND);
    ctl->stop = 1u;
  }
  for (int r = 0; r < kLists; ++r) ctl->jobs[r][0] = r == kLarge && run ? 1u : 0u;
  ctl->max_m = ctl->cut;
  ctl->pool = ctl->cut;
  ctl->work = 0;
  ctl->levels = 0;
}

// Element count of a layout launch: n, or (d_n given: a launch sized to the
// buffers' capacity n, recorded in a graph) the count the device holds.
__device__ inline int32_t live_count(int32_t n, const uint32_t* d_n) {
  if (!d_n) return n;
  const uint32_t c = *d_n;
  return c < uint32_t(n) ? int32_t(c) : n;
}

// Leaves of the output tree: one atomic per workgroup.
__global__ __launch_bounds__(256) void count_leaves_kernel(const OutNode* __restrict__ out, int32_t n,
                                                           const uint32_t* d_n, Ctl* ctl) {
  __shared__ uint32_t s_n;
  if (threadIdx.x == 0) s_n = 0;
  n = live_count(n, d_n);
  __syncthreads();
  const int32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o < n && out[o].kind == 1) atomicAdd(&s_n, 1u);
  __syncthreads();
  if (threadIdx.x == 0 && s_n) atomicAdd(&ctl->n_leaves, s_n);
}

__global__ void reset_level_kernel(Ctl* ctl) {
  for (int r = 0; r < kLists; ++r) ctl->jobs[r][0] = 0;
  ctl->pool = 0;
  ctl->max_m = 0;
  ctl->groups = 0;
  ctl->groups_small = 0;
}

// ---- layout of the traversal tree ------------------------------------------------
// Flags over all n slots (0 past the live count, so a scan over the
// capacity gives the live total).
__global__ void branch_flag_kernel(const OutNode* __restrict__ out, int32_t n, const uint32_t* d_n, uint32_t* flag) {
  const int32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t live = live_count(n, d_n);
  if (o < n) flag[o] = (o < live && out[o].kind == 0 && out[o].nchild >= 2) ? 1u : 0u;
}

__global__ void trav_nodes_kernel(const OutNode* __restrict__ out, int32_t n, const uint32_t* d_n,
                                  const uint32_t* __restrict__ flag,
                                  const uint32_t* __restrict__ tindex, BvhNode* __restrict__ nodes,
                                  const uint32_t* stop) {
  const int32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  if ((stop && *stop) || o >= live_count(n, d_n) || !flag[o]) return;
  BvhNode nd;
  float* boxes = reinterpret_cast<float*>(&nd);
  for (int c = 0; c < 2; ++c) {
    const int32_t k = out[o].child[c];
    int32_t kk = k;
    if (out[kk].kind == 0 && out[kk].nchild == 1) kk = out[kk].child[0];  // branch over one leaf
    if (out[kk].kind == 0) {
      nd.ref[c] = int32_t(tindex[kk]);
      nd.cnt[c] = 0;
    } else {
      nd.ref[c] = out[kk].prim_off;
      nd.cnt[c] = out[kk].prim_cnt;
    }
    for (int a = 0; a < 6; ++a) boxes[6 * c + a] = out[k].box[a];
  }
  nodes[tindex[o]] = nd;
}

// n-ary trees: a branch of k children becomes k - 1 binary traversal nodes,
// a balanced binary tree over its children in order (the first half, rounded
// up, on the left), numbered in preorder from the branch's base index.  The
// nodes after the first are virtual: they exist only for the traversal, and
// BVH.traversalSteps does not count them.  Walked with the reference's stack
// order (children pushed in order, the last popped first), the real nodes
// are visited in the reference's order, and a virtual node's box, the union
// of its children's, is missed only where all of them are.
__device__ inline void child_slot(const OutNode* __restrict__ out, int32_t k, const uint32_t* __restrict__ tindex,
                                  int32_t& ref, int32_t& cnt) {
  int32_t kk = k;
  if (out[kk].kind == 0 && out[kk].nchild == 1) kk = out[kk].child[0];  // branch over one leaf
  if (out[kk].kind == 0) {
    ref = int32_t(tindex[kk]);
    cnt = 0;
  } else {
    ref = out[kk].prim_off;
    cnt = out[kk].prim_cnt;
  }
}

__global__ void nary_nodes_kernel(const OutNode* __restrict__ out, int32_t n, const uint32_t* __restrict__ flag,
                                  const uint32_t* __restrict__ tindex, BvhNode* __restrict__ nodes,
                                  uint8_t* __restrict__ virt) {
  const int32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n || !flag[o]) return;
  const int32_t k = out[o].nchild, c0 = out[o].child[0];
  const uint32_t base = tindex[o];
  struct Range {
    int32_t lo, hi;
    uint32_t idx;
  };
  Range st[kMaxBranching];
  int sp = 0;
  st[sp++] = Range{0, k, base};
  while (sp > 0) {
    const Range r = st[--sp];
    const int32_t mid = r.lo + (r.hi - r.lo + 1) / 2;
    BvhNode nd;
    float* boxes = reinterpret_cast<float*>(&nd);
    const int32_t part[2][2] = {{r.lo, mid}, {mid, r.hi}};
    uint32_t next = r.idx + 1;
    for (int c = 0; c < 2; ++c) {
      const int32_t lo = part[c][0], hi = part[c][1];
      float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
      for (int32_t j = lo; j < hi; ++j)
        for (int a = 0; a < 3; ++a) {
          b[a] = fminf(b[a], out[c0 + j].box[a]);
          b[3 + a] = fmaxf(b[3 + a], out[c0 + j].box[3 + a]);
        }
      for (int a = 0; a < 6; ++a) boxes[6 * c + a] = b[a];
      if (hi - lo == 1) {
        child_slot(out, c0 + lo, tindex, nd.ref[c], nd.cnt[c]);
      } else {
        nd.ref[c] = int32_t(next);
        nd.cnt[c] = 0;
        if (sp < kMaxBranching) st[sp++] = Range{lo, hi, next};
        next += uint32_t(hi - lo - 1);
      }
    }
    nodes[r.idx] = nd;
    virt[r.idx] = r.idx == base ? 0 : 1;
  }
}

// Base index of each canonically ordered branch: the running sum of the
// traversal nodes (children - 1) of the branches before it.
__global__ void nary_weight_kernel(const OutNode* __restrict__ out, const uint32_t* __restrict__ sorted,
                                   int32_t n_real, uint32_t* __restrict__ w) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n_real) w[r] = uint32_t(out[sorted[r]].nchild - 1);
}

__global__ void nary_base_kernel(const uint32_t* __restrict__ sorted, const uint32_t* __restrict__ base,
                                 int32_t n_real, uint32_t* __restrict__ tindex) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n_real) tindex[sorted[r]] = base[r];
}

// Canonical numbering of the traversal nodes.  Jobs allocate output nodes with
// atomics, so output indices vary from run to run; the two-child branches are
// numbered instead in (level, first triangle) order, which is unique per node
// (the nodes of one level cover disjoint, non-empty triangle ranges) and puts
// the root first.  The tree's layout, and every sum taken over its nodes in
// index order, are then the same on every run.
__global__ void canon_keys_kernel(const OutNode* __restrict__ out, int32_t n, const uint32_t* d_n,
                                  const uint32_t* __restrict__ flag, const uint32_t* __restrict__ tindex,
                                  int prim_bits, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= live_count(n, d_n) || !flag[o]) return;
  const uint32_t i = tindex[o];
  keys[i] = (uint64_t(uint32_t(out[o].level)) << prim_bits) | uint64_t(uint32_t(out[o].prim_off));
  vals[i] = uint32_t(o);
}

__global__ void canon_rank_kernel(const uint32_t* __restrict__ sorted, int32_t n_trav, const uint32_t* d_n,
                                  uint32_t* __restrict__ tindex) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < live_count(n_trav, d_n)) tindex[sorted[r]] = uint32_t(r);
}

// A sort over the capacity n (recorded in a graph): the slots past the live
// branch count hold keys of all ones, which order after every real key
// (level and first triangle both below their fields' maxima).
__global__ void canon_pad_kernel(const uint32_t* __restrict__ d_n, int32_t n, uint64_t mask,
                                 uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && i >= live_count(n, d_n)) {
    keys[i] = mask;
    vals[i] = 0xffffffffu;
  }
}

// Eight lanes per output node (most leaves hold a few triangles; branches
// have nothing to copy): copy a leaf's LBVH triangle ranges into PHR order.
constexpr int kCopyLanes = 8;
__global__ __launch_bounds__(256) void leaf_copy_kernel(
    const OutNode* __restrict__ out, int32_t n_out, const int32_t* __restrict__ list, Flat f,
    const int32_t* __restrict__ perm_in, const float* __restrict__ tris_in,
    const float* __restrict__ box_in, int32_t* __restrict__ perm_out, int32_t* __restrict__ src_out,
    float* __restrict__ tris_out, float* __restrict__ box_out, int32_t n, Ctl* ctl, const uint32_t* d_n,
    const uint32_t* stop) {
  const int32_t o = int32_t((blockIdx.x * 256u + threadIdx.x) / kCopyLanes);
  const int g = int(threadIdx.x) & (kCopyLanes - 1);
  if ((stop && *stop) || o >= live_count(n_out, d_n)) return;
  const OutNode nd = out[o];
  if (nd.kind != 1) return;
  if (nd.prim_off < 0 || nd.prim_cnt < 0 || nd.prim_cnt > n - nd.prim_off) {
    if (g == 0) atomicOr(&ctl->err, 4u);
    return;
  }
  int32_t dst = nd.prim_off;
  const int32_t end = nd.prim_off + nd.prim_cnt;
  for (int32_t e = 0; e < nd.list_cnt; ++e) {
    const int32_t node = list[nd.list_off + e];
    const int32_t a = f.prim[2 * node], b = f.prim[2 * node + 1];
    if (b - a > end - dst) {  // the list must fill exactly [prim_off, end)
      if (g == 0) atomicOr(&ctl->err, 4u);
      return;
    }
    for (int32_t k = a + g; k < b; k += kCopyLanes) {
      const int32_t d = dst + (k - a);
      perm_out[d] = perm_in[k];
      src_out[d] = k;
      for (int q = 0; q < 9; ++q) tris_out[9 * size_t(d) + q] = tris_in[9 * size_t(k) + q];
      for (int q = 0; q < 6; ++q) box_out[6 * size_t(d) + q] = box_in[6 * size_t(k) + q];
    }
    dst += b - a;
  }
}

__global__ void phr_header_kernel(const OutNode* __restrict__ out, const BvhHeader* __restrict__ aux,
                                  const Ctl* __restrict__ ctl, const uint32_t* __restrict__ n_trav,
                                  int32_t n, BvhHeader* __restrict__ hdr) {
  if (ctl->stop) return;  // a replay beyond its recorded bounds keeps the previous tree
  BvhHeader h = *aux;
  const OutNode& root = out[0];
  if (root.kind == 0 && root.nchild >= 2) {
    h.root_ref = 0;  // the root is the first two-child branch in index order
    h.root_cnt = 0;
  } else {
    h.root_ref = 0;
    h.root_cnt = n;
  }
  h.n_inner = int32_t(*n_trav);
  h.n_leaves = int32_t(ctl->n_leaves);
  h.depth = int32_t(ctl->max_level);
  h.n_prims = n;
  h.phr_out = int32_t(ctl->out);
  h.phr_levels = int32_t(ctl->levels);
  *hdr = h;
}

// ---- a rebuild recorded in a graph ------------------------------------------------
// Thresholds S / 2^(alpha*d + delta) (phr.go:356-358) from the LBVH root box
// on the device: S in FP64 with the host's operation order, divided by the
// powers the host computed (one IEEE division each, so every threshold equals
// the host's).
__global__ void thresholds_kernel(const BvhHeader* __restrict__ aux, const double* __restrict__ pw,
                                  double* __restrict__ th) {
#pragma clang fp contract(off)
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= kMaxDepth) return;
  const double w = double(aux->root_hi[0]) - double(aux->root_lo[0]);
  const double hgt = double(aux->root_hi[1]) - double(aux->root_lo[1]);
  const double dep = double(aux->root_hi[2]) - double(aux->root_lo[2]);
  const double S = 2 * w * hgt + 2 * w * dep + 2 * hgt * dep;
  th[d] = S / pw[d];
}

// After the recorded levels: jobs still pending mean the frame needed more
// levels than were recorded; a counter error of the levels is a tree error.
__global__ void levels_check_kernel(Ctl* ctl, uint32_t* d_err) {
  uint32_t pending = 0;
  for (int r = 0; r < kLists; ++r) pending += ctl->cur_jobs[r];
  if (pending) {
    atomicOr(d_err, TRACE_ERR_BOUND);
    ctl->stop = 1u;
  }
  if (ctl->err) {
    atomicOr(d_err, TRACE_ERR_TREE);
    ctl->stop = 1u;
  }
}

// After the scan of the branches: more than the tree's node array holds is
// out of the recorded bounds (the layout then leaves the previous tree).
__global__ void node_cap_check_kernel(const uint32_t* __restrict__ n_trav, uint32_t cap, Ctl* ctl, uint32_t* d_err) {
  if (*n_trav > cap) {
    ctl->stop = 1u;
    atomicOr(d_err, TRACE_ERR_BOUND);
  }
}

// After the layout: the traversal kernels and the widening pass were recorded
// for trees of at most max_depth binary levels.
__global__ void layout_check_kernel(const Ctl* __restrict__ ctl, int32_t max_depth, uint32_t* d_err) {
  if (ctl->err) atomicOr(d_err, TRACE_ERR_TREE);
  if (!ctl->stop && int64_t(ctl->max_level) > int64_t(max_depth)) atomicOr(d_err, TRACE_ERR_BOUND);
}

// ---- host driver ---------------------------------------------------------------
// The level counters, read once per level through the context's pinned
// staging buffer (a direct device-to-host copy, no pageable bounce).
static_assert(sizeof(Ctl) <= kPinnedBytes, "Ctl fits the pinned staging buffer");
static int read_ctl(trace_ctx* ctx, const Ctl* d, Ctl* h) {
  HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned, d, sizeof(Ctl), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(h, ctx->h_pinned, sizeof(Ctl));
  return TRACE_OK;
}

static int clone_lbvh(trace_ctx* ctx, const trace_bvh* aux, trace_bvh* out) {
  const int64_t n = aux->n_prims;
  out->has_spheres = aux->has_spheres;
  hipStream_t st = ctx->stream;
  BvhHeader h;
  HIP_TRY(ctx, hipMemcpyAsync(&h, aux->header.p, sizeof(h), hipMemcpyDeviceToHost, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  const int64_t ninner = std::max<int32_t>(1, h.root_cnt > 0 ? 1 : h.n_inner);
  HIP_TRY(ctx, out->perm.alloc(n));
  HIP_TRY(ctx, out->tris.alloc(9 * n));
  HIP_TRY(ctx, out->prim_box.alloc(6 * n));
  HIP_TRY(ctx, out->nodes.alloc(ninner));
  HIP_TRY(ctx, out->header.alloc(1));
  HIP_TRY(ctx, out->phr_src.alloc(n));
  HIP_TRY(ctx, out->inner_range.alloc(2 * ninner));
  HIP_TRY(ctx, out->run_start.alloc(n + 1));
  HIP_TRY(ctx, hipMemcpyAsync(out->perm.p, aux->perm.p, n * 4, hipMemcpyDeviceToDevice, st));
  HIP_TRY(ctx, hipMemcpyAsync(out->tris.p, aux->tris.p, 9 * n * 4, hipMemcpyDeviceToDevice, st));
  HIP_TRY(ctx, hipMemcpyAsync(out->prim_box.p, aux->prim_box.p, 6 * n * 4, hipMemcpyDeviceToDevice, st));
  HIP_TRY(ctx, hipMemcpyAsync(out->nodes.p, aux->nodes.p, ninner * sizeof(BvhNode), hipMemcpyDeviceToDevice, st));
  HIP_TRY(ctx, hipMemcpyAsync(out->header.p, aux->header.p, sizeof(BvhHeader), hipMemcpyDeviceToDevice, st));
  HIP_TRY(ctx, hipMemcpyAsync(out->inner_range.p, aux->inner_range.p, 2 * ninner * 4, hipMemcpyDeviceToDevice, st));
  HIP_TRY(ctx, hipMemcpyAsync(out->run_start.p, aux->run_start.p, (n + 1) * 4, hipMemcpyDeviceToDevice, st));
  std::vector<int32_t> iota(n);
  for (int64_t i = 0; i < n; ++i) iota[i] = int32_t(i);
  HIP_TRY(ctx, hipMemcpyAsync(out->phr_src.p, iota.data(), n * 4, hipMemcpyHostToDevice, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  out->phr_clone = true;
  out->host_depth = h.depth;
  out->virt.release();
  return wide::widen(ctx, out);
}

// The traversal layout of the output tree: branches in canonical order as
// binary nodes (one per two-child branch; k - 1 per branch of k children,
// see nary_nodes_kernel), triangles copied into PHR leaf order, the header,
// then the 4-wide tree.
static int layout(trace_ctx* ctx, const trace_bvh* aux, int64_t n, int branching, const Flat& f, OutNode* outn,
                  const int32_t* list, Ctl* ctl, uint32_t* flag, uint32_t* tidx, uint32_t* sscr, uint32_t* n_trav,
                  int32_t* ord, double* gkey, double* rcost, uint32_t* status, int64_t pool_cap, int levels,
                  bool fixed_caps, trace_bvh* out) {
  hipStream_t st = ctx->stream;
  Ctl hc;
  int rc = read_ctl(ctx, ctl, &hc);
  if (rc) return rc;
  const int32_t n_out = int32_t(hc.out);
  out->phr_out_count = n_out;
  out->phr_levels = levels;

  branch_flag_kernel<<<int(ceil_div(n_out, 256)), 256, 0, st>>>(outn, n_out, nullptr, flag);
  count_leaves_kernel<<<int(ceil_div(n_out, 256)), 256, 0, st>>>(outn, n_out, nullptr, ctl);
  scan::exclusive_u32(st, flag, tidx, n_out, n_trav, sscr);
  uint32_t h_real = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&h_real, n_trav, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  // binary traversal nodes: one per real branch, or children - 1 when n-ary
  uint32_t h_ntrav = h_real;
  uint32_t* d_ntrav = n_trav;
  if (h_real > 1) {
    // canonical order (the buffers of the refinement loop are free by now:
    // branches number fewer than the LBVH leaves, so keys fit gkey / rcost
    // and values the order buffer)
    int prim_bits = 1, level_bits = 1;
    while ((int64_t(1) << prim_bits) < n) ++prim_bits;
    while ((int64_t(1) << level_bits) <= int64_t(hc.max_level)) ++level_bits;
    if (int64_t(h_real) > pool_cap || prim_bits + level_bits > 64)
      TRACE_FAIL(ctx, TRACE_E_DEVICE, "trace_phr_build: layout buffers too small");
    auto* keys_a = reinterpret_cast<uint64_t*>(gkey);
    auto* keys_b = reinterpret_cast<uint64_t*>(rcost);
    auto* vals_a = reinterpret_cast<uint32_t*>(ord);
    auto* vals_b = reinterpret_cast<uint32_t*>(ord + pool_cap);
    canon_keys_kernel<<<int(ceil_div(n_out, 256)), 256, 0, st>>>(outn, n_out, nullptr, flag, tidx, prim_bits, keys_a,
                                                                 vals_a);
    const int where = onesweep::sort_pairs(st, keys_a, vals_a, keys_b, vals_b, h_real, prim_bits + level_bits,
                                           status, ctx->d_err, false, nullptr, 0, onesweep::Config{});
    const uint32_t* sorted = where ? vals_b : vals_a;
    if (branching > 2) {
      // base index of each branch in canonical order (the key buffers are free again)
      auto* w = reinterpret_cast<uint32_t*>(gkey);
      auto* b = w + pool_cap;
      d_ntrav = n_trav + 1;
      nary_weight_kernel<<<int(ceil_div(h_real, 256)), 256, 0, st>>>(outn, sorted, int32_t(h_real), w);
      scan::exclusive_u32(st, w, b, h_real, d_ntrav, sscr);
      nary_base_kernel<<<int(ceil_div(h_real, 256)), 256, 0, st>>>(sorted, b, int32_t(h_real), tidx);
      HIP_TRY(ctx, hipMemcpyAsync(&h_ntrav, d_ntrav, 4, hipMemcpyDeviceToHost, st));
    } else {
      canon_rank_kernel<<<int(ceil_div(h_real, 256)), 256, 0, st>>>(sorted, int32_t(h_real), nullptr, tidx);
    }
    rc = ctx_check_device_error(ctx, "trace_phr_build");
    if (rc) return rc;
  } else if (h_real == 1 && branching > 2) {
    // the root alone: its children - 1 nodes from index 0
    OutNode root;
    HIP_TRY(ctx, hipMemcpyAsync(&root, outn, sizeof(OutNode), hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    h_ntrav = uint32_t(root.nchild - 1);
    d_ntrav = n_trav + 1;
    HIP_TRY(ctx, hipMemcpyAsync(d_ntrav, &h_ntrav, 4, hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
  }
  // fixed capacities: this tree's branches and an eighth more, for the frames
  // a captured rebuild replays (the 4-wide conversion runs over the capacity)
  HIP_TRY(ctx, out->nodes.alloc(std::max<int64_t>(
                   {int64_t(1), int64_t(h_ntrav), fixed_caps ? int64_t(h_ntrav) + int64_t(h_ntrav) / 8 + 64 : 0})));
  HIP_TRY(ctx, out->perm.alloc(n));
  HIP_TRY(ctx, out->phr_src.alloc(n));
  HIP_TRY(ctx, out->tris.alloc(9 * n));
  HIP_TRY(ctx, out->prim_box.alloc(6 * n));
  HIP_TRY(ctx, out->header.alloc(1));
  if (branching > 2) {
    HIP_TRY(ctx, out->virt.alloc(std::max<uint32_t>(1, h_ntrav)));
    nary_nodes_kernel<<<int(ceil_div(n_out, 256)), 256, 0, st>>>(outn, n_out, flag, tidx, out->nodes.p,
                                                                 out->virt.p);
  } else {
    out->virt.release();  // binary: no virtual nodes (a rebuild of a tree that was n-ary)
    trav_nodes_kernel<<<int(ceil_div(n_out, 256)), 256, 0, st>>>(outn, n_out, nullptr, flag, tidx, out->nodes.p,
                                                                 nullptr);
  }
  leaf_copy_kernel<<<int(ceil_div(int64_t(n_out) * kCopyLanes, 256)), 256, 0, st>>>(outn, n_out, list, f, aux->perm.p, aux->tris.p, aux->prim_box.p,
                                        out->perm.p, out->phr_src.p, out->tris.p, out->prim_box.p,
                                        int32_t(n), ctl, nullptr, nullptr);
  phr_header_kernel<<<1, 1, 0, st>>>(outn, aux->header.p, ctl, d_ntrav, int32_t(n), out->header.p);
  HIP_TRY(ctx, hipGetLastError());
  rc = read_ctl(ctx, ctl, &hc);
  if (rc) return rc;
  if (hc.err) TRACE_FAIL(ctx, TRACE_E_DEVICE, "trace_phr_build: layout error 0x%x", hc.err);
  // binary depth: each output level is at most ceil(log2(branching)) binary levels
  int per_level = 0;
  while ((1 << per_level) < branching) ++per_level;
  out->host_depth = int32_t(hc.max_level) * per_level;
  // fixed capacities: the traversal's depth class (stack of 128 entries up to
  // depth 63, 512 up to 255), which a captured rebuild keeps for every frame
  if (fixed_caps && out->host_depth <= 255) out->host_depth = out->host_depth <= 63 ? 63 : 255;
  return wide::widen(ctx, out);
}

// A binary rebuild recorded in a graph: the layout with every launch sized by
// the buffers' capacities, the live counts read on the device (output nodes
// from the counters, branches from the scan), the canonical order sorted over
// the most branches n triangles make with padding keys after the real ones.
static int layout_graph(trace_ctx* ctx, const trace_bvh* aux, int64_t n, const Flat& f, OutNode* outn,
                        const int32_t* list, Ctl* ctl, uint32_t* flag, uint32_t* tidx, uint32_t* sscr,
                        uint32_t* n_trav, int32_t* ord, double* gkey, double* rcost, uint32_t* status,
                        int64_t pool_cap, int32_t out_cap, int rec_levels, trace_bvh* out) {
  hipStream_t st = ctx->stream;
  const uint32_t* d_out = &ctl->out;
  const int gb = int(ceil_div(out_cap, 256));
  branch_flag_kernel<<<gb, 256, 0, st>>>(outn, out_cap, d_out, flag);
  count_leaves_kernel<<<gb, 256, 0, st>>>(outn, out_cap, d_out, ctl);
  scan::exclusive_u32(st, flag, tidx, out_cap, n_trav, sscr);
  // branches beyond the node array's capacity stop the layout
  const int64_t node_cap = std::max<int64_t>(1, std::min<int64_t>(int64_t(out->nodes.n), INT32_MAX));
  node_cap_check_kernel<<<1, 1, 0, st>>>(n_trav, uint32_t(node_cap), ctl, ctx->d_err);
  const int32_t scap = int32_t(std::min<int64_t>(node_cap, std::max<int64_t>(1, n - 1)));
  // output levels stay below the recorded job levels + 1
  int prim_bits = 1, level_bits = 1;
  while ((int64_t(1) << prim_bits) < n) ++prim_bits;
  while ((int64_t(1) << level_bits) <= int64_t(rec_levels) + 1) ++level_bits;
  const int bits = prim_bits + level_bits;
  if (int64_t(scap) > pool_cap || bits > 64)
    TRACE_FAIL(ctx, TRACE_E_DEVICE, "trace_phr_rebuild: layout buffers too small");
  auto* keys_a = reinterpret_cast<uint64_t*>(gkey);
  auto* keys_b = reinterpret_cast<uint64_t*>(rcost);
  auto* vals_a = reinterpret_cast<uint32_t*>(ord);
  auto* vals_b = reinterpret_cast<uint32_t*>(ord + pool_cap);
  canon_keys_kernel<<<gb, 256, 0, st>>>(outn, out_cap, d_out, flag, tidx, prim_bits, keys_a, vals_a);
  const uint64_t mask = bits == 64 ? ~0ull : (uint64_t(1) << bits) - 1;
  canon_pad_kernel<<<int(ceil_div(scap, 256)), 256, 0, st>>>(n_trav, scap, mask, keys_a, vals_a);
  const int where = onesweep::sort_pairs(st, keys_a, vals_a, keys_b, vals_b, scap, bits, status, ctx->d_err, false,
                                         nullptr, 0, onesweep::Config{});
  const uint32_t* sorted = where ? vals_b : vals_a;
  canon_rank_kernel<<<int(ceil_div(scap, 256)), 256, 0, st>>>(sorted, scap, n_trav, tidx);
  trav_nodes_kernel<<<gb, 256, 0, st>>>(outn, out_cap, d_out, flag, tidx, out->nodes.p, &ctl->stop);
  leaf_copy_kernel<<<int(ceil_div(int64_t(out_cap) * kCopyLanes, 256)), 256, 0, st>>>(
      outn, out_cap, list, f, aux->perm.p, aux->tris.p, aux->prim_box.p, out->perm.p, out->phr_src.p, out->tris.p,
      out->prim_box.p, int32_t(n), ctl, d_out, &ctl->stop);
  phr_header_kernel<<<1, 1, 0, st>>>(outn, aux->header.p, ctl, n_trav, int32_t(n), out->header.p);
  layout_check_kernel<<<1, 1, 0, st>>>(ctl, out->host_depth, ctx->d_err);
  HIP_TRY(ctx, hipGetLastError());
  return wide::widen(ctx, out);
}

int build(trace_ctx* ctx, const trace_bvh* aux, double alpha, double delta, int branching,
          int32_t max_cut, trace_bvh* out, Mode mode) {
#pragma clang fp contract(off)
  const int64_t n = aux->n_prims;
  hipStream_t st = ctx->stream;
  const bool graph = mode == kGraphRebuild;
  // a binary rebuild sizes its scratch for any frame of n triangles (n
  // distinct keys) and its node array for an eighth more branches than it
  // makes, so that the same rebuild recorded in a graph later finds
  // everything in place
  const bool fixed_caps = mode != kFirstBuild && branching == 2;
  if (graph && (branching != 2 || out->phr_clone || !out->phr_pw_valid || !out->phr_pw.fits(kMaxDepth)))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_rebuild: a captured rebuild needs an eager binary rebuild first");

  BvhHeader ah{};
  std::vector<double> th, pw;
  if (!graph) {
    out->n_prims = n;
    out->key_mode = -1;
    out->has_spheres = aux->has_spheres;
    out->arity = branching;
    // a rebuild starts from the state of a fresh tree
    out->phr_clone = false;
    out->phr_levels = 0;
    out->phr_out_count = 0;
    out->phr_alpha = alpha;
    out->phr_delta = delta;
    HIP_TRY(ctx, hipMemcpyAsync(&ah, aux->header.p, sizeof(ah), hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    out->phr_split_work = 0;
    if (ah.root_cnt > 0) {  // a one-leaf LBVH: the initial cut is that leaf
      out->phr_initial_cut = 1;
      return clone_lbvh(ctx, aux, out);
    }
    // thresholds S / 2^(alpha*depth + delta) (phr.go:356-358), FP64 on the host
    const float rb[6] = {ah.root_lo[0], ah.root_lo[1], ah.root_lo[2], ah.root_hi[0], ah.root_hi[1], ah.root_hi[2]};
    const double w = double(rb[3]) - double(rb[0]), hgt = double(rb[4]) - double(rb[1]),
                 dep = double(rb[5]) - double(rb[2]);
    const double S = 2 * w * hgt + 2 * w * dep + 2 * hgt * dep;
    th.resize(kMaxDepth);
    pw.resize(kMaxDepth);
    for (int d = 0; d < kMaxDepth; ++d) {
      pw[d] = std::pow(2.0, alpha * double(d) + delta);
      th[d] = S / pw[d];
    }
  }

  const int32_t U = fixed_caps ? int32_t(n) : ah.n_leaves;
  const int64_t F = 2 * int64_t(U) - 1;

  const int64_t fcap = std::min<int64_t>(2 * int64_t(std::max<int32_t>(max_cut, 1)) + 64, 2 * F + 64);
  const int64_t pool_cap = F + 64;
  const int32_t out_cap = int32_t(std::min<int64_t>(3 * n + 64, (int64_t(1) << 31) - 1));
  const int32_t list_cap = int32_t(F + 64);
  // a job of list r owns at least kMinCut[r] entries of the next pool, which
  // bounds each list by the pool's capacity
  const int32_t kMinCut[kLists] = {2, kSmallCut + 1, kGroupCut + 1, kWaveCut + 1};
  int32_t list_caps[kLists];
  for (int r = 0; r < kLists; ++r) list_caps[r] = int32_t(pool_cap / kMinCut[r] + 64);

  ScratchPlan plan;
  const size_t o_fbox = plan.take(F * 6 * sizeof(float));
  const size_t o_fch = plan.take(F * 2 * sizeof(int32_t));
  const size_t o_fsz = plan.take(F * sizeof(int32_t));
  const size_t o_fpr = plan.take(F * 2 * sizeof(int32_t));
  const size_t o_th = plan.take(kMaxDepth * sizeof(double));
  const size_t o_fa = plan.take(fcap * sizeof(int32_t));
  const size_t o_fb = plan.take(fcap * sizeof(int32_t));
  const size_t o_pa = plan.take(pool_cap * sizeof(int32_t));
  const size_t o_pb = plan.take(pool_cap * sizeof(int32_t));
  size_t o_jobs[2][kLists];
  for (int b = 0; b < 2; ++b)
    for (int r = 0; r < kLists; ++r) o_jobs[b][r] = plan.take(list_caps[r] * sizeof(Task));
  const size_t o_ord = plan.take(3 * pool_cap * sizeof(int32_t));
  const size_t o_key = plan.take(pool_cap * sizeof(double));
  const size_t o_rc = plan.take(pool_cap * sizeof(double));
  const size_t o_list = plan.take(list_cap * sizeof(int32_t));
  const size_t o_ctl = plan.take(sizeof(Ctl));
  const size_t o_flag = plan.take(out_cap * sizeof(uint32_t));
  const size_t o_tidx = plan.take(out_cap * sizeof(uint32_t));
  const size_t o_scan = plan.take(scan::scratch_words(out_cap) * sizeof(uint32_t));
  const size_t o_ntr = plan.take(64);
  // n-ary build: the groups of the current and the next step (every group's
  // cut to split holds two or more members of the pool)
  const int32_t group_cap = int32_t(pool_cap / 2 + 64);
  const size_t o_ga = plan.take(branching > 2 ? group_cap * sizeof(Group) : 0);
  const size_t o_gb = plan.take(branching > 2 ? group_cap * sizeof(Group) : 0);
  // status words of the canonical-order sort (at most n - 1 branches, 8 passes)
  const size_t o_cst = plan.take(onesweep::status_words(std::max<int64_t>(n, 1), 64, 0, onesweep::Config{}) *
                                 sizeof(uint32_t));
  int rc = ctx_reserve_scratch(ctx, plan.off);
  if (rc) return rc;
  char* base = static_cast<char*>(ctx->scratch);
  Flat f;
  f.box = reinterpret_cast<float*>(base + o_fbox);
  f.child = reinterpret_cast<int32_t*>(base + o_fch);
  f.size = reinterpret_cast<int32_t*>(base + o_fsz);
  f.prim = reinterpret_cast<int32_t*>(base + o_fpr);
  auto* d_th = reinterpret_cast<double*>(base + o_th);
  auto* fa = reinterpret_cast<int32_t*>(base + o_fa);
  auto* fb = reinterpret_cast<int32_t*>(base + o_fb);
  int32_t* pools[2] = {reinterpret_cast<int32_t*>(base + o_pa), reinterpret_cast<int32_t*>(base + o_pb)};
  Task* jobs[2][kLists];
  for (int b = 0; b < 2; ++b)
    for (int r = 0; r < kLists; ++r) jobs[b][r] = reinterpret_cast<Task*>(base + o_jobs[b][r]);
  auto* ord = reinterpret_cast<int32_t*>(base + o_ord);
  auto* gkey = reinterpret_cast<double*>(base + o_key);
  auto* rcost = reinterpret_cast<double*>(base + o_rc);
  auto* list = reinterpret_cast<int32_t*>(base + o_list);
  auto* ctl = reinterpret_cast<Ctl*>(base + o_ctl);
  auto* flag = reinterpret_cast<uint32_t*>(base + o_flag);
  auto* tidx = reinterpret_cast<uint32_t*>(base + o_tidx);
  auto* sscr = reinterpret_cast<uint32_t*>(base + o_scan);
  auto* n_trav = reinterpret_cast<uint32_t*>(base + o_ntr);
  Group* gbuf[2] = {reinterpret_cast<Group*>(base + o_ga), reinterpret_cast<Group*>(base + o_gb)};

  if (graph) {
    if (!out->phr_out.fits(size_t(out_cap) * sizeof(OutNode)) || !out->nodes.fits(1))
      TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_rebuild: a captured rebuild needs an eager binary rebuild first");
  } else {
    HIP_TRY(ctx, out->phr_out.alloc(size_t(out_cap) * sizeof(OutNode)));
  }
  OutNode* outn = reinterpret_cast<OutNode*>(out->phr_out.p);

  if (graph) {
    thresholds_kernel<<<int(ceil_div(kMaxDepth, 256)), 256, 0, st>>>(aux->header.p, out->phr_pw.p, d_th);
  } else {
    HIP_TRY(ctx, hipMemcpyAsync(d_th, th.data(), kMaxDepth * sizeof(double), hipMemcpyHostToDevice, st));
    out->phr_pw_valid = false;
    if (fixed_caps) {
      HIP_TRY(ctx, out->phr_pw.alloc(kMaxDepth));
      HIP_TRY(ctx, hipMemcpyAsync(out->phr_pw.p, pw.data(), kMaxDepth * sizeof(double), hipMemcpyHostToDevice, st));
      out->phr_pw_valid = true;
      out->phr_pw_alpha = alpha;
      out->phr_pw_delta = delta;
    }
  }
  fill_async<uint32_t>(st, reinterpret_cast<uint32_t*>(ctl), sizeof(Ctl) / 4, 0u);
  flat_kernel<<<int(ceil_div(F, 256)), 256, 0, st>>>(aux->nodes.p, aux->parent.p, aux->parent.p + n,
                                                    aux->inner_range.p, aux->run_start.p,
                                                    aux->header.p, f);
  initial_cut_kernel<<<1, 1024, 0, st>>>(f, d_th, max_cut, fa, fb, int32_t(fcap), pools[0],
                                         int32_t(pool_cap), jobs[0][kLarge], ctl, aux->header.p);
  Ctl hc{};
  if (!graph) {
    rc = read_ctl(ctx, ctl, &hc);
    if (rc) return rc;
    if (hc.err) TRACE_FAIL(ctx, TRACE_E_DEVICE, "trace_phr_build: initial cut exceeded its buffers");
    out->phr_initial_cut = hc.cut;
    if (hc.cut <= 1) return clone_lbvh(ctx, aux, out);  // the whole LBVH is attached as is
  }

  const Caps caps{pool_cap, out_cap, list_cap};
  int cur = 0;
  int levels = 0;
  if (branching > 2) {
    // one step splits one cut of every group (see nary_kernel)
    const NCaps ncaps{pool_cap, out_cap, list_cap, group_cap};
    nary_root_kernel<<<1, 1, 0, st>>>(jobs[0][kLarge], int32_t(n), gbuf[0]);
    uint32_t n_groups = 1, n_small = 0, max_m = hc.cut;  // the root's group is on the large list
    while (n_groups + n_small > 0) {
      int lds_p = 0;
      if (max_m > 64u && max_m <= uint32_t(kSortLds)) {
        lds_p = 128;
        while (uint32_t(lds_p) < max_m) lds_p <<= 1;
      }
      const size_t lds_bytes = size_t(lds_p) * (sizeof(double) + sizeof(int32_t));
      reset_level_kernel<<<1, 1, 0, st>>>(ctl);
      const int nb = int(n_groups);
      if (nb > 0) {
        int bsz = 64;
        if (max_m > 64u && nb < 2048) bsz = (nb <= 128 && max_m > 256u) ? 512 : 256;
        if (bsz == 512)
          nary_kernel<512><<<nb, 512, lds_bytes, st>>>(f, d_th, gbuf[cur], gbuf[cur ^ 1], pools[cur], pools[cur ^ 1],
                                                         ord, gkey, rcost, outn, list, ctl, ncaps, lds_p, branching);
        else if (bsz == 256)
          nary_kernel<256><<<nb, 256, lds_bytes, st>>>(f, d_th, gbuf[cur], gbuf[cur ^ 1], pools[cur], pools[cur ^ 1],
                                                         ord, gkey, rcost, outn, list, ctl, ncaps, lds_p, branching);
        else
          nary_kernel<64><<<nb, 64, lds_bytes, st>>>(f, d_th, gbuf[cur], gbuf[cur ^ 1], pools[cur], pools[cur ^ 1],
                                                       ord, gkey, rcost, outn, list, ctl, ncaps, lds_p, branching);
      }
      if (n_small > 0)
        nary_lane_kernel<<<int(ceil_div(n_small, 256)), 256, 0, st>>>(f, d_th, gbuf[cur], int32_t(n_small),
                                                                      gbuf[cur ^ 1], pools[cur], pools[cur ^ 1], outn,
                                                                      list, ctl, ncaps, branching);
      rc = read_ctl(ctx, ctl, &hc);
      if (rc) return rc;
      if (hc.err) TRACE_FAIL(ctx, TRACE_E_DEVICE, "trace_phr_build: device error 0x%x at step %d", hc.err, levels);
      if (int64_t(hc.groups) + hc.groups_small > group_cap)
        TRACE_FAIL(ctx, TRACE_E_DEVICE, "trace_phr_build: %u + %u groups exceed the %d the step buffer holds",
                   hc.groups, hc.groups_small, group_cap);
      n_groups = hc.groups;
      n_small = hc.groups_small;
      max_m = hc.max_m;
      cur ^= 1;
      ++levels;
    }
    out->phr_split_work = int64_t(hc.work);
    return layout(ctx, aux, n, branching, f, outn, list, ctl, flag, tidx, sscr, n_trav, ord, gkey, rcost,
                  reinterpret_cast<uint32_t*>(base + o_cst), pool_cap, levels, false, out);
  }
  // Device-driven levels, kLevelBatch per read of the counters: each level's
  // kernels take its job counts from the device (advance_level_kernel), so
  // the host only learns after a batch whether jobs remain; levels past the
  // last one find no jobs.  Every job holds a cut of at least two members and
  // makes one split attempt (branching 2), so BuildWithCost's counter
  // (phr.go:221) gains the level's pool size: the initial cut, then what each
  // level appends for the next.
  constexpr int kLevelBatch = 8;
  const size_t lds_full = size_t(kSortLds) * (sizeof(double) + sizeof(int32_t));
  const size_t lds_small = size_t(kSmallLds) * (sizeof(double) + sizeof(int32_t));
  const int n_cu = std::max(1, ctx->n_cu);
  auto level = [&]() {
    Lists nx;
    for (int r = 0; r < kLists; ++r) {
      nx.p[r] = jobs[cur ^ 1][r];
      nx.cap[r] = list_caps[r];
    }
    const Task* lg = jobs[cur][kLarge];
    // the large jobs: one of the three launches is the level's (sel_b)
    task_dev_kernel<512><<<2048, 512, lds_full, st>>>(f, d_th, lg, pools[cur], pools[cur ^ 1], nx, ord, gkey,
                                                      rcost, outn, list, ctl, caps, kSortLds);
    task_dev_kernel<256><<<2048, 256, lds_small, st>>>(f, d_th, lg, pools[cur], pools[cur ^ 1], nx, ord, gkey,
                                                       rcost, outn, list, ctl, caps, kSmallLds);
    task_dev_kernel<64><<<n_cu * 16, 64, lds_small, st>>>(f, d_th, lg, pools[cur], pools[cur ^ 1], nx, ord, gkey,
                                                          rcost, outn, list, ctl, caps, kSmallLds);
    group_task_kernel<kWaveCut><<<n_cu * 8, 256, 0, st>>>(f, d_th, jobs[cur][2], pools[cur], pools[cur ^ 1], nx,
                                                          outn, list, ctl, caps);
    group_task_kernel<kGroupCut><<<n_cu * 4, 256, 0, st>>>(f, d_th, jobs[cur][1], pools[cur], pools[cur ^ 1], nx,
                                                           outn, list, ctl, caps);
    small_task_kernel<<<n_cu * 2, 256, 0, st>>>(f, d_th, jobs[cur][0], pools[cur], pools[cur ^ 1], nx, outn, list,
                                                ctl, caps);
    cur ^= 1;
  };
  seed_levels_kernel<<<1, 1, 0, st>>>(ctl, graph ? ctx->d_err : nullptr);
  advance_level_kernel<<<1, 1, 0, st>>>(ctl);  // later levels are advanced by the small jobs' launch
  if (graph) {
    // the levels of the last eager rebuild and a few more; levels past the
    // frame's last one find no jobs
    const int rec = std::max(1, out->phr_levels) + kGraphSpareLevels;
    for (int k = 0; k < rec; ++k) level();
    levels_check_kernel<<<1, 1, 0, st>>>(ctl, ctx->d_err);
    HIP_TRY(ctx, hipGetLastError());
    return layout_graph(ctx, aux, n, f, outn, list, ctl, flag, tidx, sscr, n_trav, ord, gkey, rcost,
                        reinterpret_cast<uint32_t*>(base + o_cst), pool_cap, out_cap, rec, out);
  }
  for (;;) {
    for (int k = 0; k < kLevelBatch; ++k) level();
    rc = read_ctl(ctx, ctl, &hc);
    if (rc) return rc;
    if (hc.err) TRACE_FAIL(ctx, TRACE_E_DEVICE, "trace_phr_build: device error 0x%x by level %u", hc.err, hc.levels);
    // the last level of the batch has already advanced the counters: the
    // next level's jobs are the current ones now
    uint32_t pending = 0;
    for (int r = 0; r < kLists; ++r) pending += hc.cur_jobs[r];
    if (pending == 0) break;
  }
  levels = int(hc.levels);
  out->phr_split_work = int64_t(hc.work);
  return layout(ctx, aux, n, branching, f, outn, list, ctl, flag, tidx, sscr, n_trav, ord, gkey, rcost,
                reinterpret_cast<uint32_t*>(base + o_cst), pool_cap, levels, fixed_caps, out);
}

}  // namespace phr
