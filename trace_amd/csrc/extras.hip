// Batched rendering, on-device animation, HIP-graph capture and the
// autotuner (C-ABI entry points of include/trace_hip.h that have no reference
// counterpart).
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"
#include "onesweep.hpp"
#include "philox.hpp"
#include "render.hpp"
#include "traverse.hpp"

struct trace_graph {
  trace_ctx* ctx = nullptr;  // its scratch addresses are baked into the graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// p' = p + amplitude * sin(2*pi*frame/period + k.p) * u for every vertex; the
// frame number is read from device memory so a captured graph advances it.
__global__ void animate_kernel(const double* __restrict__ rest, double* __restrict__ pos,
                               int64_t nvert, double amplitude, double period, double kx,
                               double ky, double kz, double ux, double uy, double uz,
                               const uint32_t* __restrict__ d_frame) {
  const int64_t v = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (v >= nvert) return;
  const double f = double(*d_frame);
  const double* p = rest + 3 * v;
  if (v % 3 != 0 && prim_is_sphere(rest + 9 * (v / 3))) {  // a sphere moves its centre only
    pos[3 * v + 0] = p[0];
    pos[3 * v + 1] = p[1];
    pos[3 * v + 2] = p[2];
    return;
  }
  const double w = amplitude * sin(6.283185307179586 * f / period + (kx * p[0] + ky * p[1] + kz * p[2]));
  pos[3 * v + 0] = p[0] + w * ux;
  pos[3 * v + 1] = p[1] + w * uy;
  pos[3 * v + 2] = p[2] + w * uz;
}

__global__ void frame_step_kernel(uint32_t* d_frame) { *d_frame += 1u; }

__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, int64_t n,
                                                   float lr, float b1, float b2, float eps, float c1,
                                                   float c2, float lo, float hi) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / c1;  // bias corrections c1 = 1 - b1^t, c2 = 1 - b2^t
  const float vh = vi / c2;
  p[i] = fminf(fmaxf(p[i] - lr * mh / (sqrtf(vh) + eps), lo), hi);
}

void unit_from(double a, double b, double out[3]) {
  const double z = 1.0 - 2.0 * a;
  const double r = std::sqrt(std::max(0.0, 1.0 - z * z));
  const double phi = 6.283185307179586 * b;
  out[0] = r * std::cos(phi);
  out[1] = r * std::sin(phi);
  out[2] = z;
}

}  // namespace

extern "C" {

int trace_render_batched(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* forest,
                         int32_t nscene, const int32_t* h_seg_ids, const trace_camera* h_cams,
                         const trace_render_params* prm, float* d_sum, float* d_sumsq,
                         float* d_last) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh || !forest || !h_seg_ids || !h_cams || !prm || !d_sum || !d_sumsq || nscene < 1)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_batched: null argument or no scene");
  if (forest->base || forest->nseg < 2)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_batched: needs a forest (a segmented mesh's LBVH)");
  if (!mesh->has_shading) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_batched: mesh has no shading data");
  if (mesh->ntri != forest->n_prims) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_batched: mesh does not match the forest");
  if (prm->width < 2 || prm->height < 2 || prm->spp < 0 || prm->max_depth < 0 || prm->sample_offset < 0 ||
      prm->miss < TRACE_MISS_DEFAULT || prm->miss > TRACE_MISS_SUN)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_batched: bad render parameters");
  if (int64_t(prm->width) * prm->height > (int64_t(1) << 28))
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_render_batched: image too large");
  // one sample of every scene must fit a wavefront batch (int32 path ids)
  if (int64_t(prm->width) * prm->height * nscene > (int64_t(1) << 30))
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_render_batched: %d scenes of %dx%d exceed 2^30 paths per sample",
               nscene, prm->width, prm->height);
  std::vector<trace_bvh> views(nscene);
  std::vector<const trace_bvh*> trees(nscene);
  for (int32_t c = 0; c < nscene; ++c) {
    const int32_t s = h_seg_ids[c];
    if (s < 0 || s >= forest->nseg) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_batched: scene %d names segment %d", c, s);
    views[c].base = forest;
    views[c].seg = s;
    views[c].host_depth = forest->host_depth;
    trees[c] = &views[c];
  }
  if (prm->spp == 0) return TRACE_OK;
  return render::render_scenes(ctx, mesh, trees.data(), nscene, h_cams, prm, d_sum, d_sumsq, d_last);
}

int trace_mesh_animate(trace_ctx* ctx, trace_mesh* mesh, const trace_mesh* rest, double amplitude,
                       double period, uint64_t seed, uint32_t* d_frame) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh || !rest || !d_frame) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_animate: null argument");
  if (mesh->ntri != rest->ntri) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_animate: meshes differ in size");
  if (!(period > 0)) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_animate: period must be positive");
  const Philox4 r = philox4x32_10(0u, 0u, 0x616e696du, 0u, uint32_t(seed), uint32_t(seed >> 32));
  double u[3], kd[3];
  unit_from(double(u01(r.v[0])), double(u01(r.v[1])), u);
  unit_from(double(u01(r.v[2])), double(u01(r.v[3])), kd);
  const double kmag = 6.283185307179586 * 3.0;  // three waves across a unit length
  const int64_t nvert = 3 * mesh->ntri;
  animate_kernel<<<int(ceil_div(nvert, 256)), 256, 0, ctx->stream>>>(
      rest->pos.p, mesh->pos.p, nvert, amplitude, period, kmag * kd[0], kmag * kd[1], kmag * kd[2],
      u[0], u[1], u[2], d_frame);
  frame_step_kernel<<<1, 1, 0, ctx->stream>>>(d_frame);
  HIP_TRY(ctx, hipGetLastError());
  if (mesh->has_shading && mesh->geo_normals) return render::refresh_geo_normals(ctx, mesh);
  return TRACE_OK;
}

int trace_mesh_bind_params(trace_ctx* ctx, trace_mesh* mesh, const float* d_albedo,
                           const float* d_emission) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_bind_params: null mesh");
  mesh->tri_alb = d_albedo;
  mesh->tri_emi = d_emission;
  return TRACE_OK;
}

static int backward_impl(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh, const trace_camera* cam,
                         const trace_render_params* prm, const float* d_keep, const float* d_grad_image,
                         float* d_grad_albedo, float* d_grad_emission) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh || !bvh || !cam || !prm || !d_grad_image || !d_grad_albedo || !d_grad_emission)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_backward: null argument");
  if (!mesh->has_shading) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_backward: mesh has no shading data");
  if (mesh->has_specular)
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_render_backward: Reflective / Refractive materials are not differentiated");
  if (!bvh->base && bvh->nseg > 1)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_backward: differentiate a forest through its segment views");
  const int64_t tree_tris = bvh->base ? bvh->base->n_prims : bvh->n_prims;
  if (tree_tris != mesh->ntri) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_backward: bvh does not match the mesh");
  if (prm->width < 2 || prm->height < 2 || prm->spp < 1 || prm->max_depth < 0 || prm->sample_offset < 0 ||
      prm->miss < TRACE_MISS_DEFAULT || prm->miss > TRACE_MISS_SUN)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_backward: bad render parameters (spp >= 1 needed)");
  if (int64_t(prm->width) * prm->height > (int64_t(1) << 28))
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_render_backward: image too large");
  if (d_keep && reinterpret_cast<uintptr_t>(d_keep) % 16)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_backward_kept: d_keep must be 16-byte aligned");
  const trace_bvh* one[1] = {bvh};
  return render::backward_scenes(ctx, mesh, one, 1, cam, prm, d_grad_image, d_grad_albedo, d_grad_emission,
                                 reinterpret_cast<const float4*>(d_keep));
}

int trace_render_backward(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh,
                          const trace_camera* cam, const trace_render_params* prm,
                          const float* d_grad_image, float* d_grad_albedo, float* d_grad_emission) {
  return backward_impl(ctx, mesh, bvh, cam, prm, nullptr, d_grad_image, d_grad_albedo, d_grad_emission);
}

int trace_render_backward_kept(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh,
                               const trace_camera* cam, const trace_render_params* prm, const float* d_keep,
                               const float* d_grad_image, float* d_grad_albedo, float* d_grad_emission) {
  if (ctx && !d_keep) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_backward_kept: null d_keep");
  return backward_impl(ctx, mesh, bvh, cam, prm, d_keep, d_grad_image, d_grad_albedo, d_grad_emission);
}

int trace_adam_step(trace_ctx* ctx, float* d_param, const float* d_grad, float* d_m, float* d_v,
                    int64_t n, float lr, float beta1, float beta2, float eps, int32_t step, float lo,
                    float hi) {
  if (!ctx) return TRACE_E_INVALID;
  if (n < 0 || (n > 0 && (!d_param || !d_grad || !d_m || !d_v)) || step < 1 || !(lr > 0) ||
      !(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1) || !(eps > 0) || !(lo <= hi))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_adam_step: bad arguments");
  if (n == 0) return TRACE_OK;
  const float c1 = float(1.0 - std::pow(double(beta1), double(step)));
  const float c2 = float(1.0 - std::pow(double(beta2), double(step)));
  adam_kernel<<<int(ceil_div(n, 256)), 256, 0, ctx->stream>>>(d_param, d_grad, d_m, d_v, n, lr, beta1, beta2,
                                                             eps, c1, c2, lo, hi);
  HIP_TRY(ctx, hipGetLastError());
  return TRACE_OK;
}

int trace_render_segments(trace_ctx* ctx, int64_t* out) {
  if (!ctx) return TRACE_E_INVALID;
  if (!out) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_segments: null output");
  unsigned long long v = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&v, segment_counter(ctx), sizeof(v), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  *out = int64_t(v);
  return TRACE_OK;
}

int trace_render_counts(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh, const trace_camera* cam,
                        const trace_render_params* prm, int64_t* h_counts) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh || !bvh || !cam || !prm || !h_counts)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_counts: null argument");
  if (!mesh->has_shading) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_counts: mesh has no shading data");
  if (!bvh->base && bvh->nseg > 1)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_counts: render a forest through its segment views");
  const int64_t tree_tris = bvh->base ? bvh->base->n_prims : bvh->n_prims;
  if (tree_tris != mesh->ntri) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_counts: bvh and mesh differ");
  if (prm->width < 2 || prm->height < 2 || prm->spp < 1 || prm->max_depth < 0 || prm->sample_offset < 0 ||
      prm->miss < TRACE_MISS_DEFAULT || prm->miss > TRACE_MISS_SUN ||
      int64_t(prm->width) * prm->height > (int64_t(1) << 28))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_render_counts: bad render parameters");
  const size_t px = size_t(prm->width) * prm->height * 3;
  DevArray<float> img;
  DevArray<unsigned long long> cnt;
  HIP_TRY(ctx, img.alloc(2 * px));
  HIP_TRY(ctx, cnt.alloc(2));
  fill_async<unsigned long long>(ctx->stream, cnt.p, 2, 0ull);
  ctx->count_visits = cnt.p;
  int rc = render::render(ctx, mesh, bvh, cam, prm, img.p, img.p + px, nullptr);
  ctx->count_visits = nullptr;
  if (rc) return rc;
  unsigned long long v[3] = {0, 0, 0};
  HIP_TRY(ctx, hipMemcpyAsync(v, segment_counter(ctx), sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(v + 1, cnt.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < 3; ++k) h_counts[k] = int64_t(v[k]);
  return ctx_check_device_error(ctx, "trace_render_counts");
}

int trace_mesh_export(trace_ctx* ctx, const trace_mesh* mesh, double* h_pos) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh || !h_pos) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_export: null argument");
  HIP_TRY(ctx, hipMemcpyAsync(h_pos, mesh->pos.p, 9 * mesh->ntri * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return TRACE_OK;
}

int trace_graph_begin(trace_ctx* ctx) {
  if (!ctx) return TRACE_E_INVALID;
  if (ctx->stream == nullptr) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_graph_begin: the legacy stream cannot be captured");
  if (ctx->capturing) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_graph_begin: a capture is already open");
  HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
  ctx->capturing = true;
  return TRACE_OK;
}

int trace_graph_end(trace_ctx* ctx, trace_graph** out) {
  if (!ctx) return TRACE_E_INVALID;
  if (!out) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_graph_end: null output");
  *out = nullptr;
  if (!ctx->capturing) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_graph_end: no capture is open");
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
  ctx->capturing = false;
  trace_graph* g = new (std::nothrow) trace_graph();
  if (!g) {
    if (graph) (void)hipGraphDestroy(graph);
    TRACE_FAIL(ctx, TRACE_E_OOM, "trace_graph_end: host allocation failed");
  }
  g->graph = graph;
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    TRACE_FAIL(ctx, TRACE_E_HIP, "trace_graph_end: %s", hipGetErrorString(e));
  }
  g->ctx = ctx;
  ctx->live_graphs += 1;
  *out = g;
  return TRACE_OK;
}

int trace_graph_launch(trace_ctx* ctx, const trace_graph* graph) {
  if (!ctx) return TRACE_E_INVALID;
  if (!graph) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_graph_launch: null graph");
  HIP_TRY(ctx, hipGraphLaunch(graph->exec, ctx->stream));
  return TRACE_OK;
}

int trace_graph_destroy(trace_graph* graph) {
  if (!graph) return TRACE_E_INVALID;
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  if (graph->ctx) graph->ctx->live_graphs -= 1;
  delete graph;
  return TRACE_OK;
}

int trace_get_tune(trace_ctx* ctx, trace_tune* out) {
  if (!ctx) return TRACE_E_INVALID;
  if (!out) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_get_tune: null output");
  std::memset(out, 0, sizeof(*out));
  out->trav_threads = ctx->tune.trav_threads;
  out->render_threads = ctx->tune.render_threads;
  out->render_batch = ctx->tune.render_batch;
  out->tuned = ctx->tune.tuned ? 1 : 0;
  out->trav_persistent = ctx->tune.trav_persistent ? 1 : 0;
  out->render_persistent = ctx->tune.render_persistent ? 1 : 0;
  out->sort_bits = ctx->tune.sort_bits;
  out->sort_keys_per_thread = ctx->tune.sort_kpt;
  out->trav_order = ctx->tune.trav_order ? 1 : 0;
  out->render_order = ctx->tune.render_order;
  return TRACE_OK;
}

int trace_set_tune(trace_ctx* ctx, const trace_tune* t) {
  if (!ctx) return TRACE_E_INVALID;
  if (!t) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_set_tune: null argument");
  auto ok_threads = [](int v) { return v == 64 || v == 128 || v == 256; };
  if (!ok_threads(t->trav_threads) || !ok_threads(t->render_threads) || t->render_batch < 1)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_set_tune: workgroup sizes must be 64, 128 or 256");
  if (t->render_order < 0 || t->render_order > 2)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_set_tune: render_order must be 0, 1 or 2");
  if (!onesweep::valid_config(onesweep::Config{t->sort_bits, t->sort_keys_per_thread}))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_set_tune: sort configuration (%d bits, %d keys per thread) unsupported",
               t->sort_bits, t->sort_keys_per_thread);
  ctx->tune.trav_threads = t->trav_threads;
  ctx->tune.render_threads = t->render_threads;
  ctx->tune.render_batch = t->render_batch;
  ctx->tune.trav_persistent = t->trav_persistent != 0;
  ctx->tune.render_persistent = t->render_persistent != 0;
  ctx->tune.sort_bits = t->sort_bits;
  ctx->tune.sort_kpt = t->sort_keys_per_thread;
  ctx->tune.trav_order = t->trav_order != 0;
  ctx->tune.render_order = t->render_order;
  return TRACE_OK;
}

}  // extern "C"

namespace {

// Mean device time (ms) of `reps` runs of fn on the context stream, after one
// untimed run.
template <typename F>
int time_ms(trace_ctx* ctx, F fn, int reps, float* ms) {
  hipEvent_t a, b;
  HIP_TRY(ctx, hipEventCreate(&a));
  HIP_TRY(ctx, hipEventCreate(&b));
  int rc = fn();
  if (rc == TRACE_OK) {
    HIP_TRY(ctx, hipEventRecord(a, ctx->stream));
    for (int i = 0; i < reps && rc == TRACE_OK; ++i) rc = fn();
    HIP_TRY(ctx, hipEventRecord(b, ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(b));
    float t = 0;
    HIP_TRY(ctx, hipEventElapsedTime(&t, a, b));
    *ms = t / reps;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

}  // namespace

extern "C" {

int trace_autotune_traversal(trace_ctx* ctx, const trace_bvh* bvh, const float* d_rays, int64_t n,
                             float* h_ms) {
  if (!ctx) return TRACE_E_INVALID;
  if (!bvh || !d_rays || n <= 0) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_autotune_traversal: bad arguments");
  DevArray<char> out;
  HIP_TRY(ctx, out.alloc(size_t(n) * 8));
  int32_t* d_hit = reinterpret_cast<int32_t*>(out.p);
  float* d_t = reinterpret_cast<float*>(out.p + size_t(n) * 4);
  const int cand[3] = {64, 128, 256};
  int best = ctx->tune.trav_threads;
  bool best_p = ctx->tune.trav_persistent, best_o = ctx->tune.trav_order;
  float best_ms = INFINITY;
  // modes: one ray per lane; persistent lanes in input order; persistent lanes, rays grouped by origin
  for (int m = 0; m < 3; ++m) {
    for (int c = 0; c < 3; ++c) {
      ctx->tune.trav_threads = cand[c];
      ctx->tune.trav_persistent = m >= 1;
      ctx->tune.trav_order = m == 2;
      float ms = 0;
      int rc = time_ms(ctx, [&] { return traverse::closest_hit(ctx, bvh, d_rays, n, d_hit, d_t, nullptr, nullptr); }, 3, &ms);
      if (rc) {
        ctx->tune.trav_threads = best;
        ctx->tune.trav_persistent = best_p;
        ctx->tune.trav_order = best_o;
        return rc;
      }
      if (h_ms) h_ms[3 * m + c] = ms;
      if (ms < best_ms) {
        best_ms = ms;
        best = cand[c];
        best_p = m >= 1;
        best_o = m == 2;
      }
    }
  }
  ctx->tune.trav_threads = best;
  ctx->tune.trav_persistent = best_p;
  ctx->tune.trav_order = best_o;
  ctx->tune.tuned = true;
  return ctx_check_device_error(ctx, "trace_autotune_traversal");
}

int trace_autotune_build(trace_ctx* ctx, const trace_mesh* mesh, int key_mode, float* h_ms) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh || (key_mode != TRACE_KEY_REF63 && key_mode != TRACE_KEY_M30))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_autotune_build: bad arguments");
  const onesweep::Config cand[6] = {{8, 16}, {8, 8}, {8, 4}, {10, 8}, {10, 4}, {11, 4}};
  const TuneState before = ctx->tune;
  trace_bvh* bvh = nullptr;
  int rc = trace_lbvh_build(ctx, mesh, nullptr, key_mode, &bvh);
  if (rc) return rc;
  int best = -1;
  float best_ms = INFINITY;
  for (int c = 0; c < 6 && rc == TRACE_OK; ++c) {
    ctx->tune.sort_bits = cand[c].bits;
    ctx->tune.sort_kpt = cand[c].kpt;
    float ms = 0;
    rc = time_ms(ctx, [&] { return trace_lbvh_rebuild(ctx, mesh, nullptr, key_mode, bvh); }, 3, &ms);
    if (rc) break;
    if (h_ms) h_ms[c] = ms;
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  trace_bvh_destroy(bvh);
  if (rc || best < 0) {
    ctx->tune = before;
    return rc ? rc : TRACE_E_DEVICE;
  }
  ctx->tune.sort_bits = cand[best].bits;
  ctx->tune.sort_kpt = cand[best].kpt;
  ctx->tune.tuned = true;
  return ctx_check_device_error(ctx, "trace_autotune_build");
}

int trace_autotune_render(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh,
                          const trace_camera* cam, const trace_render_params* prm, float* h_ms) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh || !bvh || !cam || !prm || prm->spp < 1) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_autotune_render: bad arguments");
  const size_t px = size_t(prm->width) * prm->height * 3;
  DevArray<float> out;
  HIP_TRY(ctx, out.alloc(3 * px));
  auto run = [&] {
    return trace_render_device(ctx, mesh, bvh, cam, prm, out.p, out.p + px, out.p + 2 * px);
  };
  const int cand[3] = {64, 128, 256};
  int best = ctx->tune.render_threads;
  float best_ms = INFINITY;
  for (int c = 0; c < 3; ++c) {
    ctx->tune.render_threads = cand[c];
    float ms = 0;
    int rc = time_ms(ctx, run, 2, &ms);
    if (rc) return rc;
    if (h_ms) h_ms[c] = ms;
    if (ms < best_ms) {
      best_ms = ms;
      best = cand[c];
    }
  }
  ctx->tune.render_threads = best;
  const int64_t batches[4] = {int64_t(1) << 22, int64_t(1) << 24, int64_t(1) << 26, int64_t(1) << 27};
  int64_t best_b = ctx->tune.render_batch;
  best_ms = INFINITY;
  for (int c = 0; c < 4; ++c) {
    ctx->tune.render_batch = batches[c];
    float ms = 0;
    int rc = time_ms(ctx, run, 2, &ms);
    if (rc) return rc;
    if (h_ms) h_ms[3 + c] = ms;
    if (ms < best_ms) {
      best_ms = ms;
      best_b = batches[c];
    }
  }
  ctx->tune.render_batch = best_b;
  bool best_p = ctx->tune.render_persistent;
  best_ms = INFINITY;
  for (int c = 0; c < 2; ++c) {
    ctx->tune.render_persistent = c == 1;
    float ms = 0;
    int rc = time_ms(ctx, run, 2, &ms);
    if (rc) return rc;
    if (h_ms) h_ms[7 + c] = ms;
    if (ms < best_ms) {
      best_ms = ms;
      best_p = c == 1;
    }
  }
  ctx->tune.render_persistent = best_p;
  ctx->tune.tuned = true;
  return ctx_check_device_error(ctx, "trace_autotune_render");
}

}  // extern "C"
