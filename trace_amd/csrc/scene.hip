// Scene compile on the device: Scene.collectTracables (scene.go:95-114) with
// Mesh.Transformed / Triangle.transformed (scene.go:141-151,
// primitive.go:152-170) as one kernel over the flattened triangles.
//
// The scene graph itself is tiny and stays on the host: a depth-first walk
// (a node's own mesh, then its children in order) composes each node's
// matrix as node.transformation x parent (MultiplyMatrix, math.go:299-320)
// and the normal matrix (t^T)^-1 (math.go:164-297).  The per-triangle work --
// positions by t as points (w = 1, the w row dropped), vertex normals by the
// normal matrix, also as points, not renormalised -- runs on the device in
// FP64 with the reference's operation order, writing the mesh's positions,
// FP32 normals and material indices in the reference's triangle order.  A
// mesh whose own node matrix equals the identity is copied untransformed even
// below a transformed parent: that is what collectTracables does.
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"
#include "render.hpp"

struct trace_scene {
  int64_t ntri = 0;
  int32_t n_mesh = 0;
  std::vector<int64_t> off;  // n_mesh + 1 triangle offsets
  std::vector<int32_t> mesh_mat;
  DevArray<double> pos;      // 9 per triangle
  DevArray<double> nrm;      // 9 per triangle (empty: geometric normals)
  std::vector<float> mats;   // nmat x (kind, r, g, b)
  int32_t nmat = 0;
};

namespace scene {
namespace {

using Mat = double[16];

// MultiplyMatrix (math.go:299-320): out = a x b, each entry summed left to right.
void mat_mul(const double* a, const double* b, double* out) {
#pragma clang fp contract(off)
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c)
      out[4 * r + c] = a[4 * r] * b[c] + a[4 * r + 1] * b[4 + c] + a[4 * r + 2] * b[8 + c] +
                       a[4 * r + 3] * b[12 + c];
}

void transpose(const double* m, double* out) {
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) out[4 * c + r] = m[4 * r + c];
}

// Cofactor inverse (math.go:164-297).  Entry e of the adjugate is a signed sum
// of six triple products m[a] m[b] m[c], added left to right in the order
// listed; the determinant is row 0 against column 0 of the adjugate, and every
// entry is scaled by its reciprocal.  The term order is kept because it fixes
// the FP64 rounding.
struct Term {
  int8_t sign, a, b, c;
};
constexpr int kAdjOrder[16] = {0, 4, 8, 12, 1, 5, 9, 13, 2, 6, 10, 14, 3, 7, 11, 15};
constexpr Term kAdj[16][6] = {
    {{1, 5, 10, 15}, {-1, 5, 11, 14}, {-1, 9, 6, 15}, {1, 9, 7, 14}, {1, 13, 6, 11}, {-1, 13, 7, 10}},
    {{-1, 4, 10, 15}, {1, 4, 11, 14}, {1, 8, 6, 15}, {-1, 8, 7, 14}, {-1, 12, 6, 11}, {1, 12, 7, 10}},
    {{1, 4, 9, 15}, {-1, 4, 11, 13}, {-1, 8, 5, 15}, {1, 8, 7, 13}, {1, 12, 5, 11}, {-1, 12, 7, 9}},
    {{-1, 4, 9, 14}, {1, 4, 10, 13}, {1, 8, 5, 14}, {-1, 8, 6, 13}, {-1, 12, 5, 10}, {1, 12, 6, 9}},
    {{-1, 1, 10, 15}, {1, 1, 11, 14}, {1, 9, 2, 15}, {-1, 9, 3, 14}, {-1, 13, 2, 11}, {1, 13, 3, 10}},
    {{1, 0, 10, 15}, {-1, 0, 11, 14}, {-1, 8, 2, 15}, {1, 8, 3, 14}, {1, 12, 2, 11}, {-1, 12, 3, 10}},
    {{-1, 0, 9, 15}, {1, 0, 11, 13}, {1, 8, 1, 15}, {-1, 8, 3, 13}, {-1, 12, 1, 11}, {1, 12, 3, 9}},
    {{1, 0, 9, 14}, {-1, 0, 10, 13}, {-1, 8, 1, 14}, {1, 8, 2, 13}, {1, 12, 1, 10}, {-1, 12, 2, 9}},
    {{1, 1, 6, 15}, {-1, 1, 7, 14}, {-1, 5, 2, 15}, {1, 5, 3, 14}, {1, 13, 2, 7}, {-1, 13, 3, 6}},
    {{-1, 0, 6, 15}, {1, 0, 7, 14}, {1, 4, 2, 15}, {-1, 4, 3, 14}, {-1, 12, 2, 7}, {1, 12, 3, 6}},
    {{1, 0, 5, 15}, {-1, 0, 7, 13}, {-1, 4, 1, 15}, {1, 4, 3, 13}, {1, 12, 1, 7}, {-1, 12, 3, 5}},
    {{-1, 0, 5, 14}, {1, 0, 6, 13}, {1, 4, 1, 14}, {-1, 4, 2, 13}, {-1, 12, 1, 6}, {1, 12, 2, 5}},
    {{-1, 1, 6, 11}, {1, 1, 7, 10}, {1, 5, 2, 11}, {-1, 5, 3, 10}, {-1, 9, 2, 7}, {1, 9, 3, 6}},
    {{1, 0, 6, 11}, {-1, 0, 7, 10}, {-1, 4, 2, 11}, {1, 4, 3, 10}, {1, 8, 2, 7}, {-1, 8, 3, 6}},
    {{-1, 0, 5, 11}, {1, 0, 7, 9}, {1, 4, 1, 11}, {-1, 4, 3, 9}, {-1, 8, 1, 7}, {1, 8, 3, 5}},
    {{1, 0, 5, 10}, {-1, 0, 6, 9}, {-1, 4, 1, 10}, {1, 4, 2, 9}, {1, 8, 1, 6}, {-1, 8, 2, 5}},
};

void inverse(const double* m, double* out) {
#pragma clang fp contract(off)
  double adj[16];
  for (int k = 0; k < 16; ++k) {
    const Term* t = kAdj[k];
    // the first term's sign applies to its product: -m[a]*m[b]*m[c] = (-m[a])*m[b]*m[c]
    double s = (t[0].sign < 0 ? -m[t[0].a] : m[t[0].a]) * m[t[0].b] * m[t[0].c];
    for (int j = 1; j < 6; ++j) {
      const double p = m[t[j].a] * m[t[j].b] * m[t[j].c];
      s = t[j].sign < 0 ? s - p : s + p;
    }
    adj[kAdjOrder[k]] = s;
  }
  double det = m[0] * adj[0] + m[1] * adj[4] + m[2] * adj[8] + m[3] * adj[12];
  det = 1.0 / det;
  for (int k = 0; k < 16; ++k) out[k] = adj[k] * det;
}

bool is_identity(const double* m) {
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c)
      if (m[4 * r + c] != (r == c ? 1.0 : 0.0)) return false;
  return true;
}

// One mesh occurrence in the flattened order.
struct Inst {
  double t[16];     // composed transformation
  double tinv[16];  // (t^T)^-1
  int64_t src;      // first source triangle
  int64_t dst;      // first output triangle
  int32_t mat;
  int32_t raw;      // 1: copied untransformed
};

// Matrix4.Multiply of the point (x, y, z, 1), w row dropped (math.go:155-162).
__device__ inline void xform(const double* m, const double* v, double* o) {
#pragma clang fp contract(off)
  for (int r = 0; r < 3; ++r) o[r] = m[4 * r] * v[0] + m[4 * r + 1] * v[1] + m[4 * r + 2] * v[2] + m[4 * r + 3] * 1.0;
}

__global__ __launch_bounds__(256) void flatten_kernel(const Inst* __restrict__ inst, int32_t n_inst,
                                                      const double* __restrict__ src_pos,
                                                      const double* __restrict__ src_nrm, int64_t ntri,
                                                      double* __restrict__ pos, float* __restrict__ nrm,
                                                      int32_t* __restrict__ mat_idx) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= ntri) return;
  int32_t lo = 0, hi = n_inst;  // last instance with dst <= i
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (inst[mid].dst <= i) lo = mid;
    else hi = mid;
  }
  const Inst& in = inst[lo];
  const int64_t s = in.src + (i - in.dst);
  const double* p = src_pos + 9 * s;
  double* q = pos + 9 * i;
  if (prim_is_sphere(p)) {
    // Sphere.transformed (primitive.go:42-50): the centre as a point, the
    // radius unchanged
    for (int a = 0; a < 9; ++a) q[a] = p[a];
    if (!in.raw) xform(in.t, p, q);
    if (src_nrm)
      for (int a = 0; a < 9; ++a) nrm[9 * i + a] = 0.0f;
    mat_idx[i] = in.mat;
    return;
  }
  for (int v = 0; v < 3; ++v) {
    if (in.raw) {
      for (int a = 0; a < 3; ++a) q[3 * v + a] = p[3 * v + a];
    } else {
      xform(in.t, p + 3 * v, q + 3 * v);
    }
  }
  if (src_nrm) {
    const double* n = src_nrm + 9 * s;
    for (int v = 0; v < 3; ++v) {
      double o[3];
      if (in.raw) {
        o[0] = n[3 * v];
        o[1] = n[3 * v + 1];
        o[2] = n[3 * v + 2];
      } else {
        xform(in.tinv, n + 3 * v, o);
      }
      for (int a = 0; a < 3; ++a) nrm[9 * i + 3 * v + a] = float(o[a]);
    }
  }
  mat_idx[i] = in.mat;
}

}  // namespace
}  // namespace scene

extern "C" {

int trace_scene_create(trace_ctx* ctx, const double* h_pos, const double* h_nrm, const int64_t* h_mesh_off,
                       int32_t n_mesh, const int32_t* h_mesh_mat, const float* h_materials, int32_t nmat,
                       trace_scene** out) {
  if (!ctx) return TRACE_E_INVALID;
  if (!out || !h_pos || !h_mesh_off || !h_mesh_mat || !h_materials || n_mesh < 1 || nmat < 1)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_scene_create: bad arguments");
  *out = nullptr;
  if (h_mesh_off[0] != 0) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_scene_create: offsets must start at 0");
  for (int32_t k = 0; k < n_mesh; ++k) {
    if (h_mesh_off[k + 1] < h_mesh_off[k])
      TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_scene_create: mesh %d has a negative size", k);
    if (h_mesh_mat[k] < 0 || h_mesh_mat[k] >= nmat)
      TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_scene_create: mesh %d has material %d of %d", k, h_mesh_mat[k], nmat);
  }
  for (int32_t m = 0; m < nmat; ++m) {
    const float kind = h_materials[4 * m];
    if (kind != float(TRACE_MAT_DIFFUSE) && kind != float(TRACE_MAT_LIGHT))
      TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_scene_create: material %d kind %g (only Diffuse and Light)", m,
                 double(kind));
  }
  const int64_t ntri = h_mesh_off[n_mesh];
  if (ntri < 1 || ntri >= (int64_t(1) << 30))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_scene_create: %lld source triangles (need 1 .. 2^30 - 1)",
               static_cast<long long>(ntri));
  trace_scene* sc = new (std::nothrow) trace_scene();
  if (!sc) TRACE_FAIL(ctx, TRACE_E_OOM, "trace_scene_create: host allocation failed");
  sc->ntri = ntri;
  sc->n_mesh = n_mesh;
  sc->off.assign(h_mesh_off, h_mesh_off + n_mesh + 1);
  sc->mesh_mat.assign(h_mesh_mat, h_mesh_mat + n_mesh);
  sc->mats.assign(h_materials, h_materials + 4 * nmat);
  sc->nmat = nmat;
  hipError_t e = sc->pos.alloc(9 * ntri);
  if (e == hipSuccess) e = hipMemcpyAsync(sc->pos.p, h_pos, 9 * ntri * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && h_nrm) e = sc->nrm.alloc(9 * ntri);
  if (e == hipSuccess && h_nrm)
    e = hipMemcpyAsync(sc->nrm.p, h_nrm, 9 * ntri * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    delete sc;
    TRACE_FAIL(ctx, e == hipErrorOutOfMemory ? TRACE_E_OOM : TRACE_E_HIP, "trace_scene_create: %s",
               hipGetErrorString(e));
  }
  *out = sc;
  return TRACE_OK;
}

int trace_mesh_export_shading(trace_ctx* ctx, const trace_mesh* mesh, float* h_nrm, int32_t* h_mat_idx) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_export_shading: null mesh");
  if (!mesh->has_shading) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_mesh_export_shading: the mesh has no shading data");
  const size_t n = size_t(mesh->ntri);
  if (h_nrm) HIP_TRY(ctx, hipMemcpyAsync(h_nrm, mesh->nrm.p, 9 * n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  if (h_mat_idx)
    HIP_TRY(ctx, hipMemcpyAsync(h_mat_idx, mesh->mat_idx.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return TRACE_OK;
}

int trace_scene_destroy(trace_scene* scene) {
  if (!scene) return TRACE_E_INVALID;
  delete scene;
  return TRACE_OK;
}

int trace_scene_compile(trace_ctx* ctx, const trace_scene* sc, const int32_t* h_parent, const int32_t* h_node_mesh,
                        const double* h_matrix, int32_t n_nodes, trace_mesh** io_mesh) {
  if (!ctx) return TRACE_E_INVALID;
  if (!sc || !io_mesh || n_nodes < 0 || (n_nodes > 0 && (!h_parent || !h_node_mesh || !h_matrix)))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_scene_compile: bad arguments");
  std::vector<std::vector<int32_t>> kids(size_t(n_nodes) + 1);  // entry n_nodes: the root
  for (int32_t k = 0; k < n_nodes; ++k) {
    const int32_t p = h_parent[k];
    if (p < -1 || p >= k)
      TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_scene_compile: node %d has parent %d (need -1 <= parent < node)", k, p);
    if (h_node_mesh[k] < -1 || h_node_mesh[k] >= sc->n_mesh)
      TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_scene_compile: node %d names mesh %d of %d", k, h_node_mesh[k],
                 sc->n_mesh);
    kids[p < 0 ? n_nodes : p].push_back(k);
  }
  // depth-first: a node's own mesh, then its children in order (scene.go:100-114)
  std::vector<scene::Inst> inst;
  struct Frame {
    int32_t node;
    double t[16];
  };
  std::vector<Frame> st;
  Frame root{n_nodes, {}};
  for (int a = 0; a < 16; ++a) root.t[a] = (a % 5 == 0) ? 1.0 : 0.0;  // Scene.root: identity, no mesh
  int64_t dst = 0;
  // children are pushed in reverse so they pop in order
  for (auto it = kids[n_nodes].rbegin(); it != kids[n_nodes].rend(); ++it) {
    Frame f{*it, {}};
    scene::mat_mul(h_matrix + 16 * size_t(*it), root.t, f.t);
    st.push_back(f);
  }
  while (!st.empty()) {
    const Frame f = st.back();
    st.pop_back();
    const int32_t m = h_node_mesh[f.node];
    if (m >= 0 && sc->off[m + 1] > sc->off[m]) {
      scene::Inst in;
      std::memcpy(in.t, f.t, sizeof(in.t));
      double tt[16];
      scene::transpose(f.t, tt);
      scene::inverse(tt, in.tinv);
      in.src = sc->off[m];
      in.dst = dst;
      in.mat = sc->mesh_mat[m];
      in.raw = scene::is_identity(h_matrix + 16 * size_t(f.node)) ? 1 : 0;
      inst.push_back(in);
      dst += sc->off[m + 1] - sc->off[m];
    }
    for (auto it = kids[f.node].rbegin(); it != kids[f.node].rend(); ++it) {
      Frame c{*it, {}};
      scene::mat_mul(h_matrix + 16 * size_t(*it), f.t, c.t);
      st.push_back(c);
    }
  }
  if (dst < 1) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_scene_compile: the scene has no triangles");
  if (dst >= (int64_t(1) << 30))
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "trace_scene_compile: %lld triangles exceed 2^30 - 1",
               static_cast<long long>(dst));
  trace_mesh* mesh = *io_mesh;
  const bool fresh = mesh == nullptr;
  if (!fresh && (mesh->ntri != dst || mesh->nseg != 1))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_scene_compile: the mesh holds %lld triangles, the scene %lld",
               static_cast<long long>(mesh->ntri), static_cast<long long>(dst));
  if (fresh) {
    mesh = new (std::nothrow) trace_mesh();
    if (!mesh) TRACE_FAIL(ctx, TRACE_E_OOM, "trace_scene_compile: host allocation failed");
    mesh->ntri = dst;
  }
  hipStream_t st_ = ctx->stream;
  DevArray<scene::Inst> d_inst;
  hipError_t e = mesh->pos.alloc(9 * dst);
  if (e == hipSuccess) e = mesh->nrm.alloc(9 * dst);
  if (e == hipSuccess) e = mesh->mat_idx.alloc(dst);
  if (e == hipSuccess) e = mesh->mats.alloc(sc->nmat);
  if (e == hipSuccess) e = mesh->mat_param.alloc(sc->nmat);
  if (e == hipSuccess) e = d_inst.alloc(inst.size());
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_inst.p, inst.data(), inst.size() * sizeof(scene::Inst), hipMemcpyHostToDevice, st_);
  if (e == hipSuccess)
    e = hipMemcpyAsync(mesh->mats.p, sc->mats.data(), sc->nmat * sizeof(float4), hipMemcpyHostToDevice, st_);
  if (e == hipSuccess) fill_async<float>(st_, mesh->mat_param.p, size_t(sc->nmat), 0.0f);
  if (e == hipSuccess) {
    scene::flatten_kernel<<<int(ceil_div(dst, 256)), 256, 0, st_>>>(d_inst.p, int32_t(inst.size()), sc->pos.p,
                                                                     sc->nrm.p, dst, mesh->pos.p, mesh->nrm.p,
                                                                     mesh->mat_idx.p);
    e = hipGetLastError();
  }
  mesh->nmat = sc->nmat;
  mesh->has_shading = true;
  mesh->has_specular = false;
  mesh->geo_normals = sc->nrm.p == nullptr;
  int rc = TRACE_OK;
  if (e == hipSuccess && mesh->geo_normals) rc = render::refresh_geo_normals(ctx, mesh);
  if (e == hipSuccess && rc == TRACE_OK) e = hipStreamSynchronize(st_);  // d_inst is freed on return
  if (e != hipSuccess || rc != TRACE_OK) {
    if (fresh) delete mesh;
    if (rc != TRACE_OK) return rc;
    TRACE_FAIL(ctx, e == hipErrorOutOfMemory ? TRACE_E_OOM : TRACE_E_HIP, "trace_scene_compile: %s",
               hipGetErrorString(e));
  }
  *io_mesh = mesh;
  return ctx_check_device_error(ctx, "trace_scene_compile");
}

}  // extern "C"
