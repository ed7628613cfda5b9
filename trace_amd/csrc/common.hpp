// Shared host/device definitions of the gfx950 hot path: context, device
// buffers, the BVH node record and small wave64 helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "trace_hip.h"

// ---------------------------------------------------------------------------
// host-side error plumbing
// ---------------------------------------------------------------------------
// Launch parameters the autotuner may change (defaults are the untuned choice).
struct TuneState {
  int trav_threads = 64;                // closest-hit workgroup size: 64, 128 or 256
  int render_threads = 64;              // path-segment workgroup size: 64, 128 or 256
  // Paths per wavefront batch.  Path state is 72 B/path, so 2^27 paths hold
  // at most 9.7 GB of scratch (a 1920x1080x64-spp frame in one batch); large
  // batches keep the late bounces, where few paths remain, filling the GPU.
  int64_t render_batch = int64_t(1) << 27;
  bool trav_persistent = true;          // closest hit: persistent lanes refilled from a counter
  bool trav_order = true;               // persistent closest hit: rays grouped by origin cell first
  bool render_persistent = false;       // path segments: the same for the render's queues
  // path segments after the camera's, queue grouped by ray-origin cell: 0 never,
  // 1 single-scene batches of at least 2^22 paths, 2 every single-scene batch
  int render_order = 1;
  int sort_bits = 8;                    // LBVH key sort: radix digit width (8, 10 or 11)
  int sort_kpt = 16;                    // LBVH key sort: keys per thread (tiles of 256 * kpt)
  bool tuned = false;
};

constexpr size_t kPinnedBytes = 4096;

struct trace_ctx {
  TuneState tune;
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::string err;
  uint32_t* d_err = nullptr;  // device error word: bit 0 spin bound, bit 1 stack overflow
  // grow-only scratch arena, reused across builds (never shrinks)
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  uint32_t* d_work = nullptr;  // work counters of persistent kernels (64 words)
  void* h_pinned = nullptr;    // kPinnedBytes of pinned host memory for small per-step readbacks
  int n_cu = 256;              // compute units of the device
  // HIP graphs recorded from this context hold scratch addresses: while any is
  // alive (or a capture is open) the arena must not move
  int live_graphs = 0;
  bool capturing = false;
  // set during trace_render_counts: the path-segment kernels sum their
  // wide-node visits and triangle tests here (two 64-bit words)
  unsigned long long* count_visits = nullptr;
  // contexts of the same device (own stream, scratch and error word) on which
  // the PHR grid search evaluates candidates concurrently; made on first use,
  // destroyed with this context
  std::vector<trace_ctx*> workers;
};

#define TRACE_FAIL(ctx, code, ...)                                \
  do {                                                            \
    char _b[512];                                                 \
    snprintf(_b, sizeof(_b), __VA_ARGS__);                        \
    (ctx)->err = _b;                                              \
    return (code);                                                \
  } while (0)

#define HIP_TRY(ctx, call)                                                         \
  do {                                                                             \
    hipError_t _e = (call);                                                        \
    if (_e != hipSuccess) {                                                        \
      TRACE_FAIL(ctx, _e == hipErrorOutOfMemory ? TRACE_E_OOM : TRACE_E_HIP,       \
                 "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, \
                 __LINE__);                                                        \
    }                                                                              \
  } while (0)

// Device allocation owned by a handle.  Freed with hipFree in the destructor.
template <typename T>
struct DevArray {
  T* p = nullptr;
  size_t n = 0;
  DevArray() = default;
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  ~DevArray() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  bool fits(size_t count) const { return p && count <= n; }
  hipError_t alloc(size_t count) {
    if (count <= n && p) return hipSuccess;
    release();
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = sizeof(T);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), bytes);
    if (e == hipSuccess) n = count;
    return e;
  }
};

// Bump allocator over the context scratch arena.  Sizes are rounded up to 256 B.
struct ScratchPlan {
  size_t off = 0;
  size_t take(size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  }
};
int ctx_reserve_scratch(trace_ctx* ctx, size_t bytes);
int ctx_check_device_error(trace_ctx* ctx, const char* what);
// Takes (once per device, with `device` current) the device address of the
// fixed scratch array that trace_lbvh_rebuild_noalloc uses.
hipError_t noalloc_init_device(int device);

// Where a build takes its scratch memory from.  The default (base == nullptr)
// is the context's grow-only arena; a fixed workspace is a caller-provided
// device buffer of `bytes` that is never grown, and then the build also
// refuses to (re)allocate any output array of its tree: every call that
// would need device memory fails instead of allocating.
struct Workspace {
  char* base = nullptr;
  size_t bytes = 0;
  bool fixed() const { return base != nullptr; }
};

// Base address of `bytes` of scratch: the context arena (grown if needed) or
// the fixed workspace (TRACE_E_UNSUPPORTED when it is too small).
inline int workspace_base(trace_ctx* ctx, const Workspace* ws, size_t bytes, char** out, const char* what) {
  if (ws && ws->fixed()) {
    if (bytes > ws->bytes)
      TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "%s: needs %zu bytes of scratch, the fixed workspace holds %zu", what,
                 bytes, ws->bytes);
    *out = ws->base;
    return TRACE_OK;
  }
  const int rc = ctx_reserve_scratch(ctx, bytes);
  if (rc) return rc;
  *out = static_cast<char*>(ctx->scratch);
  return TRACE_OK;
}

// Sizes a tree's output array: allocates (grow-only) with the default
// workspace; with a fixed one the array must already hold `count` entries.
#define TRACE_ENSURE(ctx, ws, arr, count)                                                              \
  do {                                                                                                \
    const size_t _c = size_t(count);                                                                  \
    if ((ws) && (ws)->fixed()) {                                                                      \
      if (!(arr).fits(_c))                                                                            \
        TRACE_FAIL(ctx, TRACE_E_INVALID,                                                              \
                   "no-allocation build: array %s holds %zu entries, %zu needed (build the tree once " \
                   "with an allocating call first)",                                                  \
                   #arr, (arr).n, _c);                                                                \
    } else {                                                                                          \
      HIP_TRY(ctx, (arr).alloc(_c));                                                                  \
    }                                                                                                 \
  } while (0)
// Sets mesh->has_spheres from its device positions (synchronous); fails on a malformed sphere.
int mesh_scan_prims(trace_ctx* ctx, trace_mesh* mesh, const char* what);

// ---------------------------------------------------------------------------
// device-resident data
// ---------------------------------------------------------------------------
struct trace_mesh {
  int64_t ntri = 0;
  DevArray<double> pos;  // 9 doubles per triangle, caller order
  // shading data, caller order (set by trace_mesh_set_shading)
  DevArray<float> nrm;        // 9 floats per triangle: vertex normals n0, n1, n2
  DevArray<int32_t> mat_idx;  // material index per triangle
  DevArray<float4> mats;      // (kind, r, g, b) per material
  DevArray<float> mat_param;  // Reflective.Diffusion / Refractive.Ratio per material (else 0)
  int32_t nmat = 0;
  bool has_specular = false;  // a Reflective or Refractive material is in use
  bool has_shading = false;
  bool has_spheres = false;  // some primitive is a Sphere (prim_is_sphere)
  bool geo_normals = false;  // nrm derived from pos (recomputed when pos animates)
  // bound per-triangle parameters (caller-owned device memory, 3 floats per
  // triangle): Diffuse albedo and Light emission replacing the material colour
  const float* tri_alb = nullptr;
  const float* tri_emi = nullptr;
  // segmented meshes (many meshes in one buffer, built as a forest in one pass)
  int32_t nseg = 1;
  std::vector<int64_t> seg_off_h;   // nseg + 1 triangle offsets
  DevArray<int64_t> seg_off;        // same, on the device
  DevArray<int32_t> tri_seg;        // segment of each triangle
};

// Inner node of the binary traversal tree: both child boxes in one 64-byte
// record so an inner-node visit is one 64 B load.  cnt[c] == 0 means child c is
// an inner node with index ref[c]; cnt[c] > 0 means child c is a leaf holding
// triangles [ref[c], ref[c] + cnt[c]) of the sorted triangle array.
struct __align__(16) BvhNode {
  float lo0[3];
  float hi0[3];
  float lo1[3];
  float hi1[3];
  int32_t ref[2];
  int32_t cnt[2];
};
static_assert(sizeof(BvhNode) == 64, "node record must be 64 bytes");

// Small header that kernels read from device memory so that no host round
// trip is needed between the build and the traversal.
struct BvhHeader {
  float root_lo[3];
  float root_hi[3];
  int32_t root_ref;   // inner node index, or first triangle if root_cnt > 0
  int32_t root_cnt;   // > 0: the root is a leaf
  int32_t n_leaves;   // unique keys (LBVH) or leaves
  int32_t n_inner;
  int32_t depth;
  int32_t n_prims;
  double enc_lo[3];   // FP64 scene box (union of triangle boxes)
  double enc_hi[3];
  int32_t wroot;      // root of the 4-wide traversal tree (-1: the root is a leaf kept whole)
  int32_t n_wide;     // wide nodes of the traversal tree (a bound for every wide index)
  int32_t phr_out;    // PHR trees: output nodes (n-ary structure) of the last build, else unset
  int32_t phr_levels; // PHR trees: levels of the last build that held jobs, else unset
};

// Node of the 4-wide traversal tree (64 bytes, one load per visit).  Child
// boxes are stored as 8-bit offsets on a per-node grid: axis a of child c
// spans [o_a + qlo * s_a, o_a + qhi * s_a] with s_a = 2^(ebits_a - 127),
// evaluated as fmaf(q, s_a, o_a); the grid is chosen so that this box
// contains the child's FP32 box.  cnt == 0: inner child, ref = wide node
// index; cnt > 0: leaf of triangles [ref, ref + cnt) (the widening pass
// splits larger binary leaves, so every count fits 16 bits).
struct __align__(64) WideNode {
  float o[3];
  uint32_t meta;    // ebits x | y << 8 | z << 16 | (mask of used child slots) << 24
  uint32_t qlo[3];  // per axis: child c in bits 8c .. 8c + 7
  uint32_t qhi[3];
  int32_t ref[4];
  uint32_t cnt[2];  // child c: bits 16 (c & 1) .. of cnt[c >> 1]
};
static_assert(sizeof(WideNode) == 64, "wide node record must be 64 bytes");

// A tree as the reference stores it (for Cost/Size/export and as PHR input):
// node i has n_children children listed at child_list[child_off[i] ...].
// Leaves hold a sorted-position range.  Used for n-ary PHR trees; the LBVH
// keeps its binary form only.
struct trace_bvh {
  int64_t n_prims = 0;
  bool has_spheres = false;  // built from a mesh holding Sphere primitives
  int32_t key_mode = 0;  // -1 for PHR
  int32_t arity = 2;
  DevArray<uint64_t> keys;     // sorted Morton keys (LBVH only)
  DevArray<int32_t> perm;      // sorted position -> caller triangle index
  DevArray<float> tris;        // sorted order: v0, v1, v2 (9 floats)
  DevArray<float> prim_box;    // sorted order: lo.xyz hi.xyz (6 floats, outward-rounded)
  DevArray<BvhNode> nodes;     // binary tree (build, refit, export, PHR input)
  DevArray<WideNode> wnodes;   // 4-wide traversal tree derived from nodes
  DevArray<int32_t> parent;    // inner parents then leaf parents ((p << 1) | slot, -1 = root)
  DevArray<uint32_t> run_start;// leaf i covers [run_start[i], run_start[i+1])
  DevArray<int32_t> inner_range; // LBVH: leaf range [lo, hi] of each inner node
  DevArray<BvhHeader> header;
  // PHR trees: output nodes (n-ary structure for Cost/export) and the LBVH
  // sorted position of each PHR-ordered triangle
  DevArray<uint8_t> phr_out;
  int64_t phr_out_count = 0;
  DevArray<int32_t> phr_src;
  // PHR trees of branching above 2: per binary traversal node, 1 when it is
  // virtual (an inner node of a branch's binary split, not a node of the tree)
  DevArray<uint8_t> virt;
  // instanced scenes (trace_instances_build): this tree is the top level over
  // the instances' world boxes, its sorted primitives are instances
  bool instanced = false;
  int32_t n_inst = 0;
  int32_t blas_depth = 0;  // deepest bottom-level tree (sizes the second stack)
  int64_t flat_prims = 0;  // triangles of the flattened scene
  DevArray<uint8_t> inst;  // per instance: transform, tree, first flattened index
  DevArray<uint8_t> blas;  // per bottom-level tree: its arrays and header
  bool phr_clone = false;  // the PHR tree is the LBVH itself (initial cut of one node)
  double phr_alpha = 0, phr_delta = 0;
  int64_t phr_initial_cut = 0;
  int64_t phr_split_work = 0;  // sizes of the cuts handed to SweepSAH (BuildWithCost's counter)
  int32_t phr_levels = 0;
  // PHR trees rebuilt in place: 2^(alpha * d + delta) for every depth d, as the
  // host's pow computes it, for the (alpha, delta) of the last eager rebuild,
  // so that a captured rebuild can form its thresholds on the device
  DevArray<double> phr_pw;
  bool phr_pw_valid = false;
  double phr_pw_alpha = 0, phr_pw_delta = 0;
  // upper bound on inner nodes along any root-to-leaf path (sizes the
  // traversal stack); an LBVH over 64-bit keys never exceeds 63
  int32_t host_depth = 63;
  // forests (built from a segmented mesh): one header per segment
  int32_t nseg = 1;
  DevArray<BvhHeader> seg_headers;
  std::vector<int64_t> seg_off_h;
  // a segment view aliases a forest's arrays and reads its segment's header
  const trace_bvh* base = nullptr;
  int32_t seg = -1;

  const BvhNode* nodes_p() const { return base ? base->nodes.p : nodes.p; }
  const WideNode* wnodes_p() const { return base ? base->wnodes.p : wnodes.p; }
  const float* tris_p() const { return base ? base->tris.p : tris.p; }
  const int32_t* perm_p() const { return base ? base->perm.p : perm.p; }
  const BvhHeader* header_p() const { return base ? base->seg_headers.p + seg : header.p; }
  int32_t depth_bound() const { return base ? base->host_depth : host_depth; }
  const uint8_t* virt_p() const { return base ? base->virt.p : virt.p; }
  bool spheres() const { return base ? base->has_spheres : has_spheres; }
};

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
#define TRACE_ERR_SPIN 1u
#define TRACE_ERR_STACK 2u
#define TRACE_ERR_TREE 4u  // a kernel met an out-of-range node or triangle reference
// a build replayed from a graph met input beyond the bounds it was recorded
// with (more PHR levels, a deeper tree, an initial cut of one node)
#define TRACE_ERR_BOUND 8u

__device__ inline uint64_t lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ inline uint64_t ballot64(bool p) { return __ballot(p); }

// Round a double toward -inf / +inf when narrowing to float so that float
// boxes always contain the double boxes they stand for.
__device__ __host__ inline float f32_down(double x) {
  float f = (float)x;
  if ((double)f > x) f = nextafterf(f, -INFINITY);
  return f;
}
__device__ __host__ inline float f32_up(double x) {
  float f = (float)x;
  if ((double)f < x) f = nextafterf(f, INFINITY);
  return f;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

template <typename T>
__global__ void fill_kernel(T* p, size_t n, T v) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// Clears device words with a kernel (not hipMemsetAsync) so the work stays
// valid inside captured HIP graphs.
template <typename T>
inline void fill_async(hipStream_t st, T* p, size_t n, T v) {
  if (n == 0) return;
  const int64_t blocks = std::min<int64_t>(ceil_div(int64_t(n), 256), 2048);
  fill_kernel<T><<<int(blocks), 256, 0, st>>>(p, n, v);
}

// Primitives are 9 values per slot (FP64 in a mesh, FP32 in a tree's
// triangle array): a triangle's three vertices, or a Sphere (primitive.go:
// 25-86) as centre (3), radius, then NaN in the fifth value -- a tag no
// triangle carries.
template <typename T>
__host__ __device__ inline bool prim_is_sphere(const T* p) {
  return p[4] != p[4];
}

// Per-axis bounds of a primitive: the reference's Min3/Max3 of the vertices
// (util.go:47-66; the result is one of the inputs) or centre -/+ radius
// (NewSphere, primitive.go:31-40).
__host__ __device__ inline void prim_bounds(const double* p, int a, double& lo, double& hi) {
  if (prim_is_sphere(p)) {
    lo = p[a] - p[3];
    hi = p[a] + p[3];
    return;
  }
  const double x = p[a], y = p[3 + a], z = p[6 + a];
  lo = (x <= y && x <= z) ? x : ((y <= x && y <= z) ? y : z);
  hi = (x >= y && x >= z) ? x : ((y >= x && y >= z) ? y : z);
}

// Path segments traced by the context's last render or backward call: a
// 64-bit counter in its own 128-byte line of the work-counter block.
inline unsigned long long* segment_counter(trace_ctx* ctx) {
  return reinterpret_cast<unsigned long long*>(ctx->d_work + 32);
}
