// C-ABI entry points (include/trace_hip.h).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "common.hpp"
#include "lbvh.hpp"
#include "phr.hpp"
#include "render.hpp"
#include "traverse.hpp"

namespace {

bool valid_key_mode(int m) { return m == TRACE_KEY_REF63 || m == TRACE_KEY_M30; }

}  // namespace

int ctx_reserve_scratch(trace_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->scratch_bytes) return TRACE_OK;
  if (ctx->capturing || ctx->live_graphs > 0)
    TRACE_FAIL(ctx, TRACE_E_INVALID,
               "scratch arena must grow to %zu bytes while %s; run the work once before capturing",
               bytes, ctx->capturing ? "a graph capture is open" : "graphs recorded from this context are alive");
  if (ctx->scratch) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipFree(ctx->scratch));
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
  }
  const size_t want = std::max(bytes, size_t(1) << 20);
  HIP_TRY(ctx, hipMalloc(&ctx->scratch, want));
  ctx->scratch_bytes = want;
  return TRACE_OK;
}

int ctx_check_device_error(trace_ctx* ctx, const char* what) {
  uint32_t e = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&e, ctx->d_err, sizeof(e), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (e) {
    HIP_TRY(ctx, hipMemsetAsync(ctx->d_err, 0, sizeof(uint32_t), ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    TRACE_FAIL(ctx, TRACE_E_DEVICE, "%s: device reported error word 0x%x (%s%s%s%s)", what, e,
               (e & TRACE_ERR_SPIN) ? "bounded wait expired; " : "",
               (e & TRACE_ERR_STACK) ? "traversal stack overflow; " : "",
               (e & TRACE_ERR_TREE) ? "malformed tree reference; " : "",
               (e & TRACE_ERR_BOUND) ? "a replayed graph's build exceeded the bounds it was recorded with" : "");
  }
  return TRACE_OK;
}

// Bit 0: some primitive is a Sphere; bit 1: a Sphere with a radius that is not
// positive and finite, or a centre that is not finite.
__global__ void scan_prims_kernel(const double* __restrict__ pos, int64_t n, uint32_t* __restrict__ flags) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  bool sph = false, bad = false;
  if (i < n) {
    const double* p = pos + 9 * i;
    sph = prim_is_sphere(p);
    bad = sph && !(p[3] > 0.0 && p[3] < INFINITY && fabs(p[0]) < INFINITY && fabs(p[1]) < INFINITY &&
                   fabs(p[2]) < INFINITY);
  }
  const uint64_t bs = ballot64(sph), bb = ballot64(bad);
  if (__lane_id() == 0 && (bs || bb)) atomicOr(flags, (bs ? 1u : 0u) | (bb ? 2u : 0u));
}

// Sets mesh->has_spheres from the device positions; rejects malformed spheres.
int mesh_scan_prims(trace_ctx* ctx, trace_mesh* mesh, const char* what) {
  DevArray<uint32_t> flag;
  HIP_TRY(ctx, flag.alloc(1));
  fill_async<uint32_t>(ctx->stream, flag.p, 1, 0u);
  scan_prims_kernel<<<int(ceil_div(mesh->ntri, 256)), 256, 0, ctx->stream>>>(mesh->pos.p, mesh->ntri, flag.p);
  uint32_t h = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&h, flag.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (h & 2u) TRACE_FAIL(ctx, TRACE_E_INVALID, "%s: a sphere needs a finite centre and a positive finite radius", what);
  mesh->has_spheres = (h & 1u) != 0;
  return TRACE_OK;
}

extern "C" {

int trace_version(void) { return 1; }

int trace_ctx_create(int device, trace_ctx** out) {
  if (!out) return TRACE_E_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return TRACE_E_NODEVICE;
  if (device < 0 || device >= count) return TRACE_E_NODEVICE;
  trace_ctx* ctx = new (std::nothrow) trace_ctx();
  if (!ctx) return TRACE_E_OOM;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&ctx->d_err), sizeof(uint32_t)) != hipSuccess ||
      hipMemset(ctx->d_err, 0, sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&ctx->d_work), 64 * sizeof(uint32_t)) != hipSuccess ||
      hipHostMalloc(&ctx->h_pinned, kPinnedBytes, hipHostMallocDefault) != hipSuccess ||
      hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
    if (ctx->d_err) (void)hipFree(ctx->d_err);
    if (ctx->d_work) (void)hipFree(ctx->d_work);
    if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return TRACE_E_HIP;
  }
  if (ctx->n_cu < 1) ctx->n_cu = 1;
  // a failure is kept and reported by trace_lbvh_rebuild_noalloc, the only user
  (void)noalloc_init_device(device);
  ctx->stream = ctx->own_stream;
  *out = ctx;
  return TRACE_OK;
}

int trace_ctx_destroy(trace_ctx* ctx) {
  if (!ctx) return TRACE_E_INVALID;
  for (trace_ctx* w : ctx->workers) (void)trace_ctx_destroy(w);
  ctx->workers.clear();
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->d_err) (void)hipFree(ctx->d_err);
  if (ctx->d_work) (void)hipFree(ctx->d_work);
  if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return TRACE_OK;
}

const char* trace_last_error(const trace_ctx* ctx) {
  if (!ctx) return "null context";
  return ctx->err.c_str();
}

int trace_ctx_set_stream(trace_ctx* ctx, void* hip_stream) {
  if (!ctx) return TRACE_E_INVALID;
  ctx->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->own_stream;
  return TRACE_OK;
}

int trace_ctx_synchronize(trace_ctx* ctx) {
  if (!ctx) return TRACE_E_INVALID;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ctx_check_device_error(ctx, "trace_ctx_synchronize");
}

// ---- mesh ------------------------------------------------------------------
static int mesh_create_impl(trace_ctx* ctx, const double* pos, int64_t ntri, trace_mesh** out,
                            hipMemcpyKind kind) {
  if (!ctx) return TRACE_E_INVALID;
  if (!out || !pos || ntri <= 0) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_create: need positions and ntri > 0 (got %lld)", (long long)ntri);
  if (ntri >= (int64_t(1) << 30)) TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_mesh_create: ntri %lld exceeds 2^30 - 1", (long long)ntri);
  *out = nullptr;
  trace_mesh* m = new (std::nothrow) trace_mesh();
  if (!m) TRACE_FAIL(ctx, TRACE_E_OOM, "trace_mesh_create: host allocation failed");
  m->ntri = ntri;
  hipError_t e = m->pos.alloc(9 * ntri);
  if (e == hipSuccess) e = hipMemcpyAsync(m->pos.p, pos, 9 * ntri * sizeof(double), kind, ctx->stream);
  if (e == hipSuccess && kind == hipMemcpyHostToDevice) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    delete m;
    TRACE_FAIL(ctx, e == hipErrorOutOfMemory ? TRACE_E_OOM : TRACE_E_HIP, "trace_mesh_create: %s", hipGetErrorString(e));
  }
  const int rc = mesh_scan_prims(ctx, m, "trace_mesh_create");
  if (rc) {
    delete m;
    return rc;
  }
  *out = m;
  return TRACE_OK;
}

int trace_mesh_create(trace_ctx* ctx, const double* h_pos, int64_t ntri, trace_mesh** out) {
  return mesh_create_impl(ctx, h_pos, ntri, out, hipMemcpyHostToDevice);
}

int trace_mesh_create_device(trace_ctx* ctx, const double* d_pos, int64_t ntri, trace_mesh** out) {
  return mesh_create_impl(ctx, d_pos, ntri, out, hipMemcpyDeviceToDevice);
}

int trace_mesh_create_segmented(trace_ctx* ctx, const double* h_pos, int64_t ntri,
                                const int64_t* h_seg_offsets, int32_t nseg, trace_mesh** out) {
  if (!ctx) return TRACE_E_INVALID;
  if (!h_seg_offsets || nseg < 1) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_create_segmented: need nseg >= 1 offsets");
  if (h_seg_offsets[0] != 0 || h_seg_offsets[nseg] != ntri)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_create_segmented: offsets must run from 0 to ntri");
  for (int32_t s = 0; s < nseg; ++s)
    if (h_seg_offsets[s + 1] <= h_seg_offsets[s])
      TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_create_segmented: segment %d is empty or offsets decrease", s);
  int rc = mesh_create_impl(ctx, h_pos, ntri, out, hipMemcpyHostToDevice);
  if (rc) return rc;
  trace_mesh* m = *out;
  m->nseg = nseg;
  m->seg_off_h.assign(h_seg_offsets, h_seg_offsets + nseg + 1);
  std::vector<int32_t> tseg(ntri);
  for (int32_t s = 0; s < nseg; ++s)
    for (int64_t i = h_seg_offsets[s]; i < h_seg_offsets[s + 1]; ++i) tseg[i] = s;
  hipError_t e = m->tri_seg.alloc(ntri);
  if (e == hipSuccess) e = m->seg_off.alloc(nseg + 1);
  if (e == hipSuccess) e = hipMemcpyAsync(m->tri_seg.p, tseg.data(), ntri * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(m->seg_off.p, h_seg_offsets, (nseg + 1) * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    delete m;
    *out = nullptr;
    TRACE_FAIL(ctx, TRACE_E_HIP, "trace_mesh_create_segmented: %s", hipGetErrorString(e));
  }
  return TRACE_OK;
}

int trace_bvh_segment(trace_ctx* ctx, const trace_bvh* forest, int32_t seg, trace_bvh** out) {
  if (!ctx) return TRACE_E_INVALID;
  if (!forest || !out) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_segment: null argument");
  if (forest->base) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_segment: already a segment view");
  if (forest->nseg < 2) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_segment: not a forest");
  if (seg < 0 || seg >= forest->nseg) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_segment: segment %d of %d", seg, forest->nseg);
  trace_bvh* v = new (std::nothrow) trace_bvh();
  if (!v) TRACE_FAIL(ctx, TRACE_E_OOM, "trace_bvh_segment: host allocation failed");
  v->base = forest;
  v->seg = seg;
  v->n_prims = forest->seg_off_h[seg + 1] - forest->seg_off_h[seg];
  v->key_mode = forest->key_mode;
  v->arity = 2;
  v->host_depth = forest->host_depth;
  *out = v;
  return TRACE_OK;
}

int trace_mesh_update(trace_ctx* ctx, trace_mesh* mesh, const double* h_pos) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh || !h_pos) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_update: null argument");
  HIP_TRY(ctx, hipMemcpyAsync(mesh->pos.p, h_pos, 9 * mesh->ntri * sizeof(double),
                              hipMemcpyHostToDevice, ctx->stream));
  const int src = mesh_scan_prims(ctx, mesh, "trace_mesh_update");
  if (src) return src;
  if (mesh->has_shading && mesh->geo_normals) {
    const int rc = render::refresh_geo_normals(ctx, mesh);
    if (rc) return rc;
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return TRACE_OK;
}

int trace_mesh_destroy(trace_mesh* mesh) {
  if (!mesh) return TRACE_E_INVALID;
  delete mesh;
  return TRACE_OK;
}

// ---- LBVH ------------------------------------------------------------------
int trace_lbvh_build(trace_ctx* ctx, const trace_mesh* mesh, const double* h_enclosing,
                     int key_mode, trace_bvh** out) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh || !out) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_lbvh_build: null argument");
  if (!valid_key_mode(key_mode)) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_lbvh_build: unknown key mode %d", key_mode);
  *out = nullptr;
  trace_bvh* bvh = new (std::nothrow) trace_bvh();
  if (!bvh) TRACE_FAIL(ctx, TRACE_E_OOM, "trace_lbvh_build: host allocation failed");
  int rc = lbvh::build(ctx, mesh, h_enclosing, key_mode, bvh);
  if (rc == TRACE_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      ctx->err = std::string("trace_lbvh_build: ") + hipGetErrorString(e);
      rc = TRACE_E_HIP;
    } else {
      rc = ctx_check_device_error(ctx, "trace_lbvh_build");
    }
  }
  if (rc != TRACE_OK) {
    delete bvh;
    return rc;
  }
  *out = bvh;
  return TRACE_OK;
}

int trace_lbvh_rebuild(trace_ctx* ctx, const trace_mesh* mesh, const double* h_enclosing,
                       int key_mode, trace_bvh* bvh) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh || !bvh) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_lbvh_rebuild: null argument");
  if (!valid_key_mode(key_mode)) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_lbvh_rebuild: unknown key mode %d", key_mode);
  if (bvh->key_mode < 0 || bvh->base) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_lbvh_rebuild: not an LBVH handle");
  if (mesh->ntri != bvh->n_prims || mesh->nseg != bvh->nseg)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_lbvh_rebuild: mesh shape differs from the one the tree was built for");
  return lbvh::build(ctx, mesh, h_enclosing, key_mode, bvh);
}

int trace_bvh_destroy(trace_bvh* bvh) {
  if (!bvh) return TRACE_E_INVALID;
  delete bvh;
  return TRACE_OK;
}

static int read_header(trace_ctx* ctx, const trace_bvh* bvh, BvhHeader* h) {
  HIP_TRY(ctx, hipMemcpyAsync(h, bvh->header_p(), sizeof(BvhHeader), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return TRACE_OK;
}

int trace_bvh_get_info(trace_ctx* ctx, const trace_bvh* bvh, trace_bvh_info* out) {
  if (!ctx) return TRACE_E_INVALID;
  if (!bvh || !out) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_get_info: null argument");
  BvhHeader h;
  int rc = read_header(ctx, bvh, &h);
  if (rc) return rc;
  std::memset(out, 0, sizeof(*out));
  out->n_prims = bvh->n_prims;
  out->n_leaves = h.n_leaves;
  // a view's header holds forest totals as bounds; its own tree has one inner
  // node fewer than leaves
  out->n_inner = h.root_cnt > 0 ? 0 : (bvh->base ? h.n_leaves - 1 : h.n_inner);
  out->n_segments = bvh->base ? 1 : bvh->nseg;
  out->depth = h.depth;
  out->key_mode = bvh->key_mode;
  out->arity = bvh->arity;
  for (int a = 0; a < 3; ++a) {
    out->root_box[a] = h.root_lo[a];
    out->root_box[3 + a] = h.root_hi[a];
  }
  if (bvh->key_mode < 0) {
    // the tree's own levels from its header (a graph replay updates them);
    // the build statistics are the last eager build's
    int per_level = 0;
    while ((1 << per_level) < bvh->arity) ++per_level;
    out->depth = bvh->phr_clone ? h.depth : h.depth * per_level;
    out->phr_initial_cut = bvh->phr_initial_cut;
    out->phr_levels = bvh->phr_clone ? 0 : bvh->arity == 2 ? h.phr_levels : bvh->phr_levels;
    out->phr_build_cost = bvh->phr_initial_cut + bvh->phr_split_work;
  }
  return TRACE_OK;
}

// Canonical preorder of a PHR tree from its n-ary output nodes: a branch's
// range is its subtree's triangle range, a leaf's its own; one-child branches
// (a cut that no split improved, phr.go:121-126) are kept, as Cost() sees them.
static int export_phr(trace_ctx* ctx, const trace_bvh* bvh, int64_t cap, int32_t* h_rec,
                      float* h_box, int64_t* n_out) {
  BvhHeader h;
  int rc = read_header(ctx, bvh, &h);
  if (rc) return rc;
  // the header's count is the last build's, a graph replay's included
  const int64_t cap_out = int64_t(bvh->phr_out.n / sizeof(phr::OutNode));
  if (h.phr_out < 1 || h.phr_out > cap_out)
    TRACE_FAIL(ctx, TRACE_E_DEVICE, "trace_bvh_export_canonical: PHR output count %d outside [1, %lld]", h.phr_out,
               static_cast<long long>(cap_out));
  std::vector<phr::OutNode> out(h.phr_out);
  HIP_TRY(ctx, hipMemcpyAsync(out.data(), bvh->phr_out.p, out.size() * sizeof(phr::OutNode),
                              hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<int32_t> st{0};
  int64_t k = 0;
  std::vector<int32_t> order;
  order.reserve(out.size());
  while (!st.empty()) {
    const int32_t o = st.back();
    st.pop_back();
    if (o < 0 || o >= int32_t(out.size()))
      TRACE_FAIL(ctx, TRACE_E_DEVICE, "trace_bvh_export_canonical: PHR node %d out of range", o);
    order.push_back(o);
    const phr::OutNode& nd = out[o];
    if (nd.kind == 0)  // a branch's children are consecutive output nodes
      for (int c = nd.nchild - 1; c >= 0; --c) st.push_back(nd.child[0] + c);
  }
  *n_out = int64_t(order.size());
  if (!h_rec || !h_box) return TRACE_OK;
  if (cap < int64_t(order.size()))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_export_canonical: capacity %lld < %lld nodes",
               (long long)cap, (long long)order.size());
  for (int32_t o : order) {
    const phr::OutNode& nd = out[o];
    int32_t* r = h_rec + 4 * k;
    r[0] = nd.kind == 0 ? 0 : 1;
    r[1] = nd.prim_off;
    r[2] = nd.prim_off + nd.prim_cnt - 1;
    r[3] = nd.kind == 0 ? nd.nchild : 0;
    for (int a = 0; a < 6; ++a) h_box[6 * k + a] = nd.box[a];
    ++k;
  }
  return TRACE_OK;
}

int trace_phr_build(trace_ctx* ctx, const trace_bvh* aux, double alpha, double delta,
                    int branching, int32_t max_cut, trace_bvh** out) {
  if (!ctx) return TRACE_E_INVALID;
  if (!aux || !out) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_build: null argument");
  *out = nullptr;
  if (aux->key_mode < 0) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_build: the auxiliary tree must be an LBVH");
  if (aux->instanced) TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_phr_build: instanced trees are not refined");
  if (aux->base || aux->nseg > 1) TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_phr_build: forests and segment views are not refined by this build");
  if (branching < 2 || branching > 8)
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_phr_build: branching factor %d (this build supports 2 to 8)", branching);
  if (!(alpha >= 0) || !std::isfinite(delta) || max_cut < 1)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_build: need alpha >= 0, finite delta, max_cut >= 1");
  trace_bvh* b = new (std::nothrow) trace_bvh();
  if (!b) TRACE_FAIL(ctx, TRACE_E_OOM, "trace_phr_build: host allocation failed");
  int rc = phr::build(ctx, aux, alpha, delta, branching, max_cut, b, phr::kFirstBuild);
  if (rc == TRACE_OK) rc = ctx_check_device_error(ctx, "trace_phr_build");
  if (rc != TRACE_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return TRACE_OK;
}

int trace_phr_rebuild(trace_ctx* ctx, const trace_bvh* aux, double alpha, double delta, int branching,
                      int32_t max_cut, trace_bvh* phr) {
  if (!ctx) return TRACE_E_INVALID;
  if (!aux || !phr) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_rebuild: null argument");
  if (phr == aux || phr->key_mode >= 0 || phr->instanced || phr->base || phr->nseg > 1)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_rebuild: the target must be a PHR tree made by trace_phr_build");
  if (aux->key_mode < 0) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_rebuild: the auxiliary tree must be an LBVH");
  if (aux->instanced) TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_phr_rebuild: instanced trees are not refined");
  if (aux->base || aux->nseg > 1)
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_phr_rebuild: forests and segment views are not refined by this build");
  if (branching < 2 || branching > 8)
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_phr_rebuild: branching factor %d (this build supports 2 to 8)",
               branching);
  if (!(alpha >= 0) || !std::isfinite(delta) || max_cut < 1)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_rebuild: need alpha >= 0, finite delta, max_cut >= 1");
  if (ctx->capturing) {
    // recorded in a graph: the launches of an eager binary rebuild of the same
    // settings, with every count read on the device (no host round trip)
    if (branching != 2)
      TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_phr_rebuild: a rebuild under graph capture takes branching 2, not %d",
                 branching);
    if (phr->arity != 2 || phr->phr_clone || !phr->phr_pw_valid || phr->phr_pw_alpha != alpha ||
        phr->phr_pw_delta != delta || phr->n_prims != aux->n_prims)
      TRACE_FAIL(ctx, TRACE_E_INVALID,
                 "trace_phr_rebuild: run this rebuild once eagerly (branching 2, the same alpha and delta, an "
                 "auxiliary tree of %lld triangles, an initial cut of two or more nodes) before capturing it",
                 static_cast<long long>(aux->n_prims));
    return phr::build(ctx, aux, alpha, delta, branching, max_cut, phr, phr::kGraphRebuild);
  }
  int rc = phr::build(ctx, aux, alpha, delta, branching, max_cut, phr, phr::kRebuild);
  if (rc == TRACE_OK) rc = ctx_check_device_error(ctx, "trace_phr_rebuild");
  return rc;
}

int trace_phr_build_with_cost(trace_ctx* ctx, const trace_bvh* aux, double alpha, double delta,
                              int branching, int32_t max_cut, trace_bvh** out, int64_t* build_cost) {
  if (!ctx) return TRACE_E_INVALID;
  if (!build_cost) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_build_with_cost: null cost output");
  int rc = trace_phr_build(ctx, aux, alpha, delta, branching, max_cut, out);
  if (rc) return rc;
  *build_cost = (*out)->phr_initial_cut + (*out)->phr_split_work;
  return TRACE_OK;
}

// Worker contexts for concurrent evaluations (at most kGridWorkers, one
// hardware queue each on the default four-queue configuration).
constexpr int kGridWorkers = 4;

static int ensure_workers(trace_ctx* ctx, int n) {
  while (int(ctx->workers.size()) < n) {
    trace_ctx* w = nullptr;
    const int rc = trace_ctx_create(ctx->device, &w);
    if (rc) TRACE_FAIL(ctx, rc, "trace_phr_optimize_grid: could not create a worker context (%d)", rc);
    w->tune = ctx->tune;
    ctx->workers.push_back(w);
  }
  for (int i = 0; i < n; ++i) ctx->workers[i]->tune = ctx->tune;
  return TRACE_OK;
}

int trace_phr_optimize_grid(trace_ctx* ctx, const trace_bvh* aux, const double* h_alphas, int32_t n_alpha,
                            const double* h_deltas, int32_t n_delta, int branching, int32_t max_cut,
                            int32_t frame_width, int32_t frame_height, double* h_sah, int64_t* h_build,
                            double* alpha_out, double* delta_out) {
  if (!ctx) return TRACE_E_INVALID;
  if (!aux || !h_alphas || !h_deltas || n_alpha < 1 || n_delta < 1 || frame_width < 1 || frame_height < 1 ||
      !alpha_out || !delta_out)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_optimize_grid: bad arguments");
  const int64_t n_eval = int64_t(n_alpha) * n_delta;
  std::vector<double> sah(n_eval);
  std::vector<int64_t> build(n_eval);
  // every (alpha, delta), alphas outer (optimizer.go:47-61).  Evaluation k
  // goes to worker k mod W, which rebuilds each of its candidates into the
  // storage of its first: a PHR build's upper levels hold few jobs and its
  // host reads counters between level groups, so W builds in flight on W
  // streams fill the GPU where one does not.  Each result is the same as a
  // sequential evaluation's (every build is deterministic).
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // aux is complete before other streams read it
  const int nw = int(std::min<int64_t>(kGridWorkers, n_eval));
  int rc = ensure_workers(ctx, nw);
  if (rc) return rc;
  std::vector<int> wrc(nw, TRACE_OK);
  auto evaluate = [&](int w) {
    trace_ctx* c = ctx->workers[w];
    if (hipSetDevice(c->device) != hipSuccess) {
      c->err = "trace_phr_optimize_grid: hipSetDevice failed in a worker thread";
      wrc[w] = TRACE_E_HIP;
      return;
    }
    trace_bvh* t = nullptr;
    for (int64_t k = w; k < n_eval; k += nw) {
      const double a = h_alphas[k / n_delta], d = h_deltas[k % n_delta];
      int r = t ? trace_phr_rebuild(c, aux, a, d, branching, max_cut, t)
                : trace_phr_build(c, aux, a, d, branching, max_cut, &t);
      if (r == TRACE_OK) {
        build[k] = t->phr_initial_cut + t->phr_split_work;
        r = trace_bvh_cost(c, t, &sah[k]);
      }
      if (r) {
        wrc[w] = r;
        break;
      }
    }
    if (t) trace_bvh_destroy(t);
  };
  std::vector<std::thread> threads;
  int started = 0;
  try {
    for (; started < nw; ++started) threads.emplace_back(evaluate, started);
  } catch (...) {
    // no thread to spare: the rest run on this one
  }
  for (int w = started; w < nw; ++w) evaluate(w);
  for (auto& th : threads) th.join();
  for (int w = 0; w < nw; ++w)
    if (wrc[w]) {
      ctx->err = ctx->workers[w]->err;
      return wrc[w];
    }
  // scores (optimizer.go:63-94): build / max build + omega^2 * SAH / max SAH,
  // omega = (prims / PRIM_THRESH)^2 * pixels / RES_THRESH; the first minimum wins
  int64_t max_build = 0;
  double max_sah = 0.0;
  for (int64_t k = 0; k < n_eval; ++k) {
    max_build = std::max(max_build, build[k]);
    max_sah = std::max(max_sah, sah[k]);
  }
  const double r = double(aux->n_prims) / 2000000.0;
  const double omega = (r * r) * (double(int64_t(frame_width) * frame_height) / 200000.0);
  double best = DBL_MAX;
  int64_t best_k = 0;
  for (int64_t k = 0; k < n_eval; ++k) {
    const double c = double(build[k]) / double(max_build) + (omega * omega) * (sah[k] / max_sah);
    if (c < best) {
      best = c;
      best_k = k;
    }
  }
  *alpha_out = h_alphas[best_k / n_delta];
  *delta_out = h_deltas[best_k % n_delta];
  if (h_sah) std::copy(sah.begin(), sah.end(), h_sah);
  if (h_build) std::copy(build.begin(), build.end(), h_build);
  return TRACE_OK;
}

int trace_phr_export_source(trace_ctx* ctx, const trace_bvh* phr, int32_t* h_src) {
  if (!ctx) return TRACE_E_INVALID;
  if (!phr || !h_src) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_export_source: null argument");
  if (phr->key_mode >= 0) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_export_source: not a PHR tree");
  HIP_TRY(ctx, hipMemcpyAsync(h_src, phr->phr_src.p, phr->n_prims * sizeof(int32_t),
                              hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return TRACE_OK;
}

int trace_bvh_export_sorted(trace_ctx* ctx, const trace_bvh* bvh, uint64_t* h_keys, int32_t* h_perm) {
  if (!ctx) return TRACE_E_INVALID;
  if (!bvh) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_export_sorted: null bvh");
  // a segment view's triangles occupy its segment's range of the forest order
  const int64_t off = bvh->base ? bvh->base->seg_off_h[bvh->seg] : 0;
  const uint64_t* keys = bvh->base ? bvh->base->keys.p : bvh->keys.p;
  if (h_keys) {
    if (!keys) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_export_sorted: tree has no Morton keys");
    HIP_TRY(ctx, hipMemcpyAsync(h_keys, keys + off, bvh->n_prims * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  }
  if (h_perm)
    HIP_TRY(ctx, hipMemcpyAsync(h_perm, bvh->perm_p() + off, bvh->n_prims * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (h_perm && off)  // triangle indices local to the segment
    for (int64_t i = 0; i < bvh->n_prims; ++i) h_perm[i] -= int32_t(off);
  return TRACE_OK;
}

// Host copy of the binary traversal tree.
struct HostTree {
  BvhHeader h;
  std::vector<BvhNode> nodes;
};

static int fetch_tree(trace_ctx* ctx, const trace_bvh* bvh, HostTree* t) {
  int rc = read_header(ctx, bvh, &t->h);
  if (rc) return rc;
  const int64_t ninner = t->h.root_cnt > 0 ? 0 : t->h.n_inner;
  t->nodes.resize(ninner);
  if (ninner > 0) {
    HIP_TRY(ctx, hipMemcpyAsync(t->nodes.data(), bvh->nodes_p(), ninner * sizeof(BvhNode), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  return TRACE_OK;
}

namespace {
struct Item {
  int32_t ref, cnt;  // cnt > 0: leaf
  float box[6];
};
}  // namespace

int trace_bvh_export_canonical(trace_ctx* ctx, const trace_bvh* bvh, int64_t cap, int32_t* h_rec,
                               float* h_box, int64_t* n_out) {
  if (!ctx) return TRACE_E_INVALID;
  if (!bvh || !n_out) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_export_canonical: null argument");
  if (bvh->key_mode < 0 && !bvh->phr_clone) return export_phr(ctx, bvh, cap, h_rec, h_box, n_out);
  if (!bvh->base && bvh->nseg > 1)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_export_canonical: a forest is exported through its segment views");
  HostTree t;
  int rc = fetch_tree(ctx, bvh, &t);
  if (rc) return rc;
  // leaf ranges of a segment view are reported relative to its segment
  const int32_t off = bvh->base ? int32_t(bvh->base->seg_off_h[bvh->seg]) : 0;
  // Preorder with children[0] first, collected first (the node count of a
  // view is only known by walking it), inner ranges filled afterwards.
  std::vector<int32_t> rec;
  std::vector<float> box;
  std::vector<Item> stack;
  Item root;
  root.ref = t.h.root_ref;
  root.cnt = t.h.root_cnt;
  for (int a = 0; a < 3; ++a) {
    root.box[a] = t.h.root_lo[a];
    root.box[3 + a] = t.h.root_hi[a];
  }
  stack.push_back(root);
  int64_t k = 0;
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    rec.resize(4 * (k + 1));
    box.resize(6 * (k + 1));
    int32_t* r = rec.data() + 4 * k;
    std::memcpy(box.data() + 6 * k, it.box, sizeof(it.box));
    if (it.cnt > 0) {
      r[0] = 1;
      r[1] = it.ref - off;
      r[2] = it.ref + it.cnt - 1 - off;
      r[3] = 0;
    } else {
      if (it.ref < 0 || it.ref >= int32_t(t.nodes.size()))
        TRACE_FAIL(ctx, TRACE_E_DEVICE, "trace_bvh_export_canonical: node %d out of range", it.ref);
      const BvhNode& nd = t.nodes[it.ref];
      r[0] = 0;
      r[1] = -1;
      r[2] = -1;
      r[3] = 2;
      Item c0, c1;
      c0.ref = nd.ref[0];
      c0.cnt = nd.cnt[0];
      c1.ref = nd.ref[1];
      c1.cnt = nd.cnt[1];
      for (int a = 0; a < 3; ++a) {
        c0.box[a] = nd.lo0[a];
        c0.box[3 + a] = nd.hi0[a];
        c1.box[a] = nd.lo1[a];
        c1.box[3 + a] = nd.hi1[a];
      }
      stack.push_back(c1);
      stack.push_back(c0);
    }
    ++k;
  }
  const int64_t total = k;
  *n_out = total;
  if (!h_rec || !h_box) return TRACE_OK;  // size query
  if (cap < total) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_export_canonical: capacity %lld < %lld nodes", (long long)cap, (long long)total);
  std::memcpy(h_rec, rec.data(), rec.size() * sizeof(int32_t));
  std::memcpy(h_box, box.data(), box.size() * sizeof(float));
  // Fill inner ranges bottom-up: in preorder, an inner node's range is from the
  // first leaf after it to the last leaf of its subtree.  Walk backwards with a
  // stack of subtree spans.
  std::vector<std::pair<int32_t, int32_t>> spans;  // (first, last) of completed subtrees
  for (int64_t i = total - 1; i >= 0; --i) {
    int32_t* r = h_rec + 4 * i;
    if (r[0] == 1) {
      spans.push_back({r[1], r[2]});
    } else {
      auto c0 = spans.back();
      spans.pop_back();
      auto c1 = spans.back();
      spans.pop_back();
      r[1] = c0.first;
      r[2] = c1.second;
      spans.push_back({r[1], r[2]});
    }
  }
  return TRACE_OK;
}

// bounding.go:45-47, FP64 from FP32 corners
__host__ __device__ inline double surface(double lx, double ly, double lz, double hx, double hy, double hz) {
  const double w = hx - lx, h = hy - ly, d = hz - lz;
  return 2 * w * h + 2 * w * d + 2 * h * d;
}
__host__ __device__ inline double surface_f(const float* b) { return surface(b[0], b[1], b[2], b[3], b[4], b[5]); }
static double surface(const float* b) { return surface_f(b); }

// Cost() telescoped: with p(c) = SA(c) / SA(parent), a node's weight in the
// root's cost is the product of p along its path, SA(node) / SA(root).  So
// cost(root) = 2 + (sum over non-root nodes of w(node) SA(node)) / SA(root),
// w = TRAVERSAL_COST for inner nodes, INTERSECTION_COST * prims for leaves.
// (A zero-area node's children have zero area too, so the reference's
// skipping of them changes nothing.)  Each thread adds the child terms of one
// binary node; blocks reduce in LDS, then one block sums the partials in
// order: the same result on every run.
__global__ __launch_bounds__(256) void cost_terms_kernel(const BvhNode* __restrict__ nodes,
                                                         const BvhHeader* __restrict__ hdr,
                                                         double* __restrict__ partial) {
  __shared__ double s[256];
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  double v = 0.0;
  if (i < hdr->n_inner) {
    const float* b = reinterpret_cast<const float*>(nodes + i);
    for (int c = 0; c < 2; ++c) {
      const int32_t cnt = nodes[i].cnt[c];
      v += (cnt > 0 ? double(cnt) : 2.0) * surface_f(b + 6 * c);
    }
  }
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void cost_sum_kernel(const double* __restrict__ partial, int64_t n,
                                                       double* __restrict__ out) {
  __shared__ double s[256];
  double v = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += 256) v += partial[k];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

static int device_cost(trace_ctx* ctx, const trace_bvh* bvh, double* cost) {
  BvhHeader h;
  int rc = read_header(ctx, bvh, &h);
  if (rc) return rc;
  if (h.root_cnt > 0) {  // a leaf root
    *cost = double(h.root_cnt);
    return TRACE_OK;
  }
  const double sa_root = surface(h.root_lo[0], h.root_lo[1], h.root_lo[2], h.root_hi[0], h.root_hi[1],
                                 h.root_hi[2]);
  if (sa_root == 0) {
    *cost = 2.0;
    return TRACE_OK;
  }
  const int64_t blocks = std::max<int64_t>(1, ceil_div(int64_t(h.n_inner), 256));
  ScratchPlan plan;
  const size_t o_part = plan.take(blocks * sizeof(double));
  const size_t o_out = plan.take(sizeof(double));
  rc = ctx_reserve_scratch(ctx, plan.off);
  if (rc) return rc;
  char* base = static_cast<char*>(ctx->scratch);
  auto* partial = reinterpret_cast<double*>(base + o_part);
  auto* d_out = reinterpret_cast<double*>(base + o_out);
  cost_terms_kernel<<<int(blocks), 256, 0, ctx->stream>>>(bvh->nodes.p, bvh->header.p, partial);
  cost_sum_kernel<<<1, 256, 0, ctx->stream>>>(partial, blocks, d_out);
  double sum = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&sum, d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  *cost = 2.0 + sum / sa_root;
  return ctx_check_device_error(ctx, "trace_bvh_cost");
}

int trace_bvh_cost(trace_ctx* ctx, const trace_bvh* bvh, double* cost) {
  if (!ctx) return TRACE_E_INVALID;
  if (!bvh || !cost) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_cost: null argument");
  // a whole binary tree reduces on the device; a forest's tree view walks its
  // own nodes on the host
  if (!bvh->base && bvh->nseg == 1 && bvh->arity == 2) return device_cost(ctx, bvh, cost);
  int64_t n = 0;
  int rc = trace_bvh_export_canonical(ctx, bvh, 0, nullptr, nullptr, &n);
  if (rc) return rc;
  std::vector<int32_t> rec(4 * n);
  std::vector<float> box(6 * n);
  rc = trace_bvh_export_canonical(ctx, bvh, n, rec.data(), box.data(), &n);
  if (rc) return rc;
  // cost(leaf) = INTERSECTION_COST * prims; cost(inner) = TRAVERSAL_COST +
  // sum_c SA(c)/SA(node) * cost(c), children skipped when SA(node) == 0
  // (bvh.go:226-239).  Evaluated in reverse preorder: the children of an
  // inner record are the top entries of the stack of finished subtrees,
  // children[0] on top.
  struct Done {
    double cost, sa;
  };
  std::vector<Done> st;
  for (int64_t i = n - 1; i >= 0; --i) {
    const int32_t* r = rec.data() + 4 * i;
    const double sa = surface(box.data() + 6 * i);
    if (r[0] == 1) {
      st.push_back({1.0 * double(r[2] - r[1] + 1), sa});
    } else {
      double acc = 0.0;
      for (int c = 0; c < r[3]; ++c) {
        const Done d = st.back();
        st.pop_back();
        if (sa != 0) acc += d.sa / sa * d.cost;
      }
      st.push_back({2.0 + acc, sa});
    }
  }
  *cost = st.empty() ? 0.0 : st.back().cost;
  return TRACE_OK;
}

int trace_bvh_size(trace_ctx* ctx, const trace_bvh* bvh, int64_t* size) {
  if (!ctx) return TRACE_E_INVALID;
  if (!bvh || !size) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_bvh_size: null argument");
  return trace_bvh_export_canonical(ctx, bvh, 0, nullptr, nullptr, size);
}

// ---- closest hit -------------------------------------------------------------
int trace_closest_hit_device(trace_ctx* ctx, const trace_bvh* bvh, const float* d_rays,
                             int64_t nrays, int32_t* d_hit, float* d_t) {
  if (!ctx) return TRACE_E_INVALID;
  if (!bvh || nrays < 0 || (nrays > 0 && (!d_rays || !d_hit || !d_t)))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_closest_hit_device: bad arguments");
  return traverse::closest_hit(ctx, bvh, d_rays, nrays, d_hit, d_t, nullptr, nullptr);
}

static int closest_hit_host(trace_ctx* ctx, const trace_bvh* bvh, const float* h_rays,
                            int64_t nrays, int32_t* h_hit, float* h_t, int32_t* h_nodes,
                            int32_t* h_tris, const char* what) {
  if (!ctx) return TRACE_E_INVALID;
  if (!bvh || nrays < 0 || (nrays > 0 && (!h_rays || !h_hit || !h_t)))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "%s: bad arguments", what);
  if (nrays == 0) return TRACE_OK;
  const bool counting = h_nodes != nullptr;
  DevArray<char> buf;
  const size_t ray_b = size_t(nrays) * 8 * sizeof(float);
  const size_t out_b = size_t(nrays) * sizeof(int32_t);
  HIP_TRY(ctx, buf.alloc(ray_b + out_b * (counting ? 4 : 2)));
  float* d_rays = reinterpret_cast<float*>(buf.p);
  int32_t* d_hit = reinterpret_cast<int32_t*>(buf.p + ray_b);
  float* d_t = reinterpret_cast<float*>(buf.p + ray_b + out_b);
  int32_t* d_nodes = counting ? reinterpret_cast<int32_t*>(buf.p + ray_b + 2 * out_b) : nullptr;
  int32_t* d_tris = counting ? reinterpret_cast<int32_t*>(buf.p + ray_b + 3 * out_b) : nullptr;
  HIP_TRY(ctx, hipMemcpyAsync(d_rays, h_rays, ray_b, hipMemcpyHostToDevice, ctx->stream));
  int rc = traverse::closest_hit(ctx, bvh, d_rays, nrays, d_hit, d_t, d_nodes, d_tris);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(h_hit, d_hit, out_b, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(h_t, d_t, out_b, hipMemcpyDeviceToHost, ctx->stream));
  if (counting) {
    HIP_TRY(ctx, hipMemcpyAsync(h_nodes, d_nodes, out_b, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(h_tris, d_tris, out_b, hipMemcpyDeviceToHost, ctx->stream));
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ctx_check_device_error(ctx, what);
}

int trace_closest_hit(trace_ctx* ctx, const trace_bvh* bvh, const float* h_rays, int64_t nrays,
                      int32_t* h_hit, float* h_t) {
  return closest_hit_host(ctx, bvh, h_rays, nrays, h_hit, h_t, nullptr, nullptr,
                          "trace_closest_hit");
}

int trace_closest_hit_counts(trace_ctx* ctx, const trace_bvh* bvh, const float* h_rays,
                             int64_t nrays, int32_t* h_nodes, int32_t* h_tris) {
  if (!ctx) return TRACE_E_INVALID;
  if (!h_nodes || !h_tris) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_closest_hit_counts: null output");
  std::vector<int32_t> hit(nrays > 0 ? nrays : 0);
  std::vector<float> t(nrays > 0 ? nrays : 0);
  return closest_hit_host(ctx, bvh, h_rays, nrays, hit.data(), t.data(), h_nodes, h_tris,
                          "trace_closest_hit_counts");
}

}  // extern "C"

// ---- path tracing --------------------------------------------------------------
extern "C" {

int trace_mesh_set_shading(trace_ctx* ctx, trace_mesh* mesh, const double* h_normals,
                           const int32_t* h_mat_index, const float* h_materials, int32_t nmat) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh || !h_mat_index || !h_materials || nmat <= 0)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_set_shading: need material indices and nmat > 0");
  for (int64_t i = 0; i < mesh->ntri; ++i)
    if (h_mat_index[i] < 0 || h_mat_index[i] >= nmat)
      TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_set_shading: triangle %lld has material %d of %d",
                 (long long)i, h_mat_index[i], nmat);
  for (int32_t m = 0; m < nmat; ++m) {
    const float kind = h_materials[4 * m];
    if (kind != float(TRACE_MAT_DIFFUSE) && kind != float(TRACE_MAT_LIGHT))
      TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_mesh_set_shading: material %d kind %g (only Diffuse and Light)",
                 m, double(kind));
  }
  return render::set_shading(ctx, mesh, h_normals, h_mat_index, h_materials, nmat, nullptr);
}

int trace_mesh_set_shading_ex(trace_ctx* ctx, trace_mesh* mesh, const double* h_normals,
                              const int32_t* h_mat_index, const float* h_materials8, int32_t nmat) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh || !h_mat_index || !h_materials8 || nmat <= 0)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_set_shading_ex: need material indices and nmat > 0");
  for (int64_t i = 0; i < mesh->ntri; ++i)
    if (h_mat_index[i] < 0 || h_mat_index[i] >= nmat)
      TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_set_shading_ex: triangle %lld has material %d of %d",
                 (long long)i, h_mat_index[i], nmat);
  std::vector<float> mats(4 * size_t(nmat)), params(nmat);
  for (int32_t m = 0; m < nmat; ++m) {
    const float* e = h_materials8 + 8 * size_t(m);
    const float kind = e[0];
    if (kind != float(TRACE_MAT_DIFFUSE) && kind != float(TRACE_MAT_LIGHT) && kind != float(TRACE_MAT_REFLECTIVE) &&
        kind != float(TRACE_MAT_REFRACTIVE))
      TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_mesh_set_shading_ex: material %d kind %g", m, double(kind));
    if (kind == float(TRACE_MAT_REFLECTIVE) && !(e[4] >= 0.0f))
      TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_set_shading_ex: material %d: Diffusion %g < 0", m, double(e[4]));
    if (kind == float(TRACE_MAT_REFRACTIVE) && !(e[4] > 0.0f))
      TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_set_shading_ex: material %d: Ratio %g <= 0", m, double(e[4]));
    for (int c = 0; c < 4; ++c) mats[4 * size_t(m) + c] = e[c];
    params[m] = (kind == float(TRACE_MAT_REFLECTIVE) || kind == float(TRACE_MAT_REFRACTIVE)) ? e[4] : 0.0f;
  }
  return render::set_shading(ctx, mesh, h_normals, h_mat_index, mats.data(), nmat, params.data());
}

static int check_render_args(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh,
                             const trace_camera* cam, const trace_render_params* prm,
                             const char* what) {
  if (!mesh || !bvh || !cam || !prm) TRACE_FAIL(ctx, TRACE_E_INVALID, "%s: null argument", what);
  if (!mesh->has_shading) TRACE_FAIL(ctx, TRACE_E_INVALID, "%s: mesh has no shading data (trace_mesh_set_shading)", what);
  if (!bvh->base && bvh->nseg > 1) TRACE_FAIL(ctx, TRACE_E_INVALID, "%s: render a forest through its segment views", what);
  const int64_t tree_tris = bvh->base ? bvh->base->n_prims : bvh->n_prims;
  if (tree_tris != mesh->ntri) TRACE_FAIL(ctx, TRACE_E_INVALID, "%s: bvh has %lld triangles, mesh %lld", what, (long long)tree_tris, (long long)mesh->ntri);
  if (prm->width < 2 || prm->height < 2 || prm->spp < 0 || prm->max_depth < 0 || prm->sample_offset < 0)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "%s: need width, height >= 2 (samples divide by W-1, H-1), spp, max_depth >= 0", what);
  if (prm->miss < TRACE_MISS_DEFAULT || prm->miss > TRACE_MISS_SUN) TRACE_FAIL(ctx, TRACE_E_INVALID, "%s: unknown miss shader %d", what, prm->miss);
  if (int64_t(prm->width) * prm->height > (int64_t(1) << 28)) TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "%s: image too large", what);
  return TRACE_OK;
}

int trace_render_device(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh,
                        const trace_camera* cam, const trace_render_params* prm, float* d_sum,
                        float* d_sumsq, float* d_last) {
  if (!ctx) return TRACE_E_INVALID;
  int rc = check_render_args(ctx, mesh, bvh, cam, prm, "trace_render_device");
  if (rc) return rc;
  if (!d_sum || !d_sumsq) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_device: null output");
  if (prm->spp == 0) return TRACE_OK;
  return render::render(ctx, mesh, bvh, cam, prm, d_sum, d_sumsq, d_last);
}

int trace_render_keep(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh, const trace_camera* cam,
                      const trace_render_params* prm, float* d_sum, float* d_sumsq, float* d_last,
                      float* d_keep) {
  if (!ctx) return TRACE_E_INVALID;
  int rc = check_render_args(ctx, mesh, bvh, cam, prm, "trace_render_keep");
  if (rc) return rc;
  if (!d_sum || !d_sumsq || !d_keep) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_keep: null output");
  if (reinterpret_cast<uintptr_t>(d_keep) % 16)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_keep: d_keep must be 16-byte aligned");
  if (mesh->has_specular)
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_render_keep: Reflective / Refractive materials are not differentiated");
  if (prm->spp == 0) return TRACE_OK;
  const trace_bvh* one[1] = {bvh};
  return render::render_scenes(ctx, mesh, one, 1, cam, prm, d_sum, d_sumsq, d_last, reinterpret_cast<float4*>(d_keep));
}

int trace_render(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh,
                 const trace_camera* cam, const trace_render_params* prm, float* h_sum,
                 float* h_sumsq, float* h_last) {
  if (!ctx) return TRACE_E_INVALID;
  int rc = check_render_args(ctx, mesh, bvh, cam, prm, "trace_render");
  if (rc) return rc;
  if (!h_sum || !h_sumsq) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render: null output");
  const size_t bytes = size_t(prm->width) * prm->height * 3 * sizeof(float);
  if (prm->spp == 0) {
    std::memset(h_sum, 0, bytes);
    std::memset(h_sumsq, 0, bytes);
    if (h_last) std::memset(h_last, 0, bytes);
    return TRACE_OK;
  }
  DevArray<float> out;
  HIP_TRY(ctx, out.alloc(3 * bytes / sizeof(float)));
  float* d_sum = out.p;
  float* d_sq = out.p + bytes / sizeof(float);
  float* d_last = out.p + 2 * bytes / sizeof(float);
  rc = render::render(ctx, mesh, bvh, cam, prm, d_sum, d_sq, h_last ? d_last : nullptr);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(h_sum, d_sum, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(h_sumsq, d_sq, bytes, hipMemcpyDeviceToHost, ctx->stream));
  if (h_last) HIP_TRY(ctx, hipMemcpyAsync(h_last, d_last, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ctx_check_device_error(ctx, "trace_render");
}

}  // extern "C"
