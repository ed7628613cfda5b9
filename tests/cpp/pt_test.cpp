// Tests of the C++ pkg/pt API (include/trace_pt.hpp), written the way the
// reference's own benchmark test uses pkg/pt (pkg/benchmark/benchmark_test.go:
// tracables from a scene -> DefaultLBVH / NewPHRBuilder(...).Build() /
// BuildFromAuxilary, NewDefaultGridOptimizer(...).OptimizedPHRparams,
// NewFrameBufferAR + NewDefaultCamera + SetTransformation +
// NewBenchmarkRenderer(...).RenderToBuffer), with every device result checked
// against the FP64 oracle (oracle/oracle.cpp, linked as the checker only).
//
//   pt_test --cpu [filter]   host-only cases (math, camera, buffers, OBJ, no-device error)
//   pt_test --gpu [filter]   every case (needs a GPU)
//
// Each case prints its name before it runs; the exit status is the number of
// failed cases (0: all passed).
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <limits>
#include <string>
#include <vector>

#include "trace_pt.hpp"

// ---- the oracle's C entry points used as the checker ----------------------------------
extern "C" {
struct oracle_tree;
struct oracle_phr;
oracle_tree* oracle_lbvh_build(const double* pos, int64_t n, const double* enclosing6, int key_mode);
void oracle_tree_free(oracle_tree* t);
int64_t oracle_tree_node_count(const oracle_tree* t);
void oracle_tree_canonical(const oracle_tree* t, int32_t* rec, double* box);
double oracle_tree_cost(const oracle_tree* t);
void oracle_closest_hit(const oracle_tree* t, const float* rays, int64_t nrays, int32_t* hit, double* tout,
                        int threads);
oracle_phr* oracle_phr_build(const oracle_tree* aux, double alpha, double delta, int branching, int64_t max_cut);
void oracle_phr_free(oracle_phr* t);
double oracle_phr_cost(const oracle_phr* t);
int64_t oracle_phr_split_work(const oracle_phr* t);
void oracle_phr_closest_hit(const oracle_phr* t, const float* rays, int64_t nrays, int32_t* hit, double* tout,
                            int threads);
int64_t oracle_scene_flatten(const double* pos, const double* nrm, const int64_t* off, const int32_t* mesh_mat,
                             const int32_t* parent, const int32_t* node_mesh, const double* matrix, int32_t n_nodes,
                             double* out_pos, double* out_nrm, int32_t* out_mat);
void oracle_render_ex(const oracle_tree* t, const double* nrm, const int32_t* mat_idx, const float* mats8,
                      int32_t nmat, const double* cam, int32_t W, int32_t H, int32_t spp, int32_t max_depth,
                      int32_t miss, uint64_t seed, int threads, double* sum, double* sumsq, double* last);
void oracle_heatmap(const oracle_tree* t, const double* cam, int32_t W, int32_t H, int32_t threshold, float* rgb,
                    int32_t* counts, int threads);
void oracle_traversal_steps(const oracle_tree* t, const double* rays, int64_t nrays, int32_t* counts, int threads);
}

namespace {

constexpr int kThreads = 16;

struct Failure {
  std::string msg;
};

#define EXPECT(cond)                                                                                  \
  do {                                                                                                \
    if (!(cond)) throw Failure{std::string(__FILE__) + ":" + std::to_string(__LINE__) + ": " + #cond}; \
  } while (0)

#define EXPECT_MSG(cond, ...)                                                                      \
  do {                                                                                             \
    if (!(cond)) {                                                                                 \
      char buf_[512];                                                                              \
      std::snprintf(buf_, sizeof(buf_), __VA_ARGS__);                                              \
      throw Failure{std::string(__FILE__) + ":" + std::to_string(__LINE__) + ": " + #cond + " (" + \
                    buf_ + ")"};                                                                   \
    }                                                                                              \
  } while (0)

// Expects `fn` to throw pt::Error with the given code.
void expect_error(int code, const std::function<void()>& fn, const char* what) {
  try {
    fn();
  } catch (const pt::Error& e) {
    EXPECT_MSG(e.Code() == code, "%s: code %d, want %d (%s)", what, e.Code(), code, e.what());
    return;
  }
  throw Failure{std::string(what) + ": no pt::Error thrown"};
}

// splitmix64: the seeded inputs
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    const double u = std::max(uniform(), 1e-300), v = uniform();
    return std::sqrt(-2 * std::log(u)) * std::cos(2 * M_PI * v);
  }
};

// Small triangles scattered over the unit cube, one Diffuse material.  Corners
// are FP32-exact, so the device's FP32 boxes equal the oracle's FP64 ones.
pt::Tracables random_soup(int n, uint64_t seed) {
  Rng r{seed};
  pt::Tracables t;
  t.Materials.push_back(pt::Diffuse{pt::NewColor(0.7, 0.7, 0.7)});
  auto f32 = [](double x) { return double(float(x)); };
  for (int i = 0; i < n; ++i) {
    const pt::Vector3 c{r.uniform(), r.uniform(), r.uniform()};
    pt::Vector3 v[3];
    for (auto& p : v)
      p = pt::Vector3{f32(c.X + 0.02 * (r.uniform() - 0.5)), f32(c.Y + 0.02 * (r.uniform() - 0.5)),
                      f32(c.Z + 0.02 * (r.uniform() - 0.5))};
    t.Prims.push_back(pt::NewTriangleWithoutNormals(v[0], v[1], v[2]));
    t.MaterialIndex.push_back(0);
  }
  return t;
}

std::vector<double> positions(const pt::Tracables& t) {
  std::vector<double> pos;
  for (const auto& p : t.Prims)
    for (const auto& v : p.Vertices) pos.insert(pos.end(), {v.Position.X, v.Position.Y, v.Position.Z});
  return pos;
}

// Origins in the unit cube, directions uniform on the sphere.
std::vector<pt::Ray> incoherent_rays(int n, uint64_t seed) {
  Rng r{seed};
  std::vector<pt::Ray> rays(n);
  for (auto& q : rays) {
    q.Origin = pt::Vector3{r.uniform(), r.uniform(), r.uniform()};
    q.Direction = pt::Unit(pt::Vector3{r.normal(), r.normal(), r.normal()});
  }
  return rays;
}

std::vector<float> ray_floats(const std::vector<pt::Ray>& rays, float tmin, float tmax) {
  std::vector<float> out;
  for (const auto& q : rays)
    out.insert(out.end(), {float(q.Origin.X), float(q.Origin.Y), float(q.Origin.Z), tmin, float(q.Direction.X),
                           float(q.Direction.Y), float(q.Direction.Z), tmax});
  return out;
}

void check_hits(const std::vector<pt::Hit>& got, const std::vector<int32_t>& ohit, const std::vector<double>& ot) {
  int64_t same = 0;
  for (size_t i = 0; i < got.size(); ++i) {
    if (got[i].Index != ohit[i]) continue;
    ++same;
    if (ohit[i] >= 0)
      EXPECT_MSG(std::fabs(got[i].T - ot[i]) <= 1e-5 * std::fabs(ot[i]) + 1e-7, "ray %zu: t %.9g vs %.9g", i,
                 got[i].T, ot[i]);
    EXPECT(got[i].Ok == (ohit[i] >= 0));
  }
  const double agree = double(same) / double(got.size());
  EXPECT_MSG(agree >= 0.9999, "hit agreement %.6f", agree);
}

// ---- host-only cases ----------------------------------------------------------------------

void case_math() {
  // closed forms of the matrix builders and their product (math.go, transformation.go)
  const pt::Matrix4 ts = pt::MultiplyMatrix(pt::Translate(1, 2, 3), pt::Scale(2, 3, 4));
  const pt::Matrix4 want{2, 0, 0, 1, 0, 3, 0, 2, 0, 0, 4, 3, 0, 0, 0, 1};
  EXPECT(ts == want);
  // the product is not commutative: scale after translation scales the offset
  const pt::Matrix4 st = pt::MultiplyMatrix(pt::Scale(2, 3, 4), pt::Translate(1, 2, 3));
  EXPECT(st[3] == 2 && st[7] == 6 && st[11] == 12);
  EXPECT(pt::Rotate(pt::NewVector3(0, 1, 0), 0.0) == pt::IdentityMatrix());
  EXPECT(pt::ScaleUniform(2.5) == pt::Scale(2.5, 2.5, 2.5));
  // a quarter turn about z maps x to y and y to -x; about x maps y to z
  const pt::Matrix4 rz = pt::Rotate(pt::NewVector3(0, 0, 5), M_PI / 2);
  EXPECT(std::fabs(rz[0]) < 1e-15 && std::fabs(rz[4] - 1) < 1e-15 && std::fabs(rz[1] + 1) < 1e-15);
  const pt::Matrix4 rx = pt::Rotate(pt::NewVector3(1, 0, 0), M_PI / 2);
  EXPECT(std::fabs(rx[9] - 1) < 1e-15 && std::fabs(rx[6] + 1) < 1e-15);
  // rotations are orthonormal: R R^T = I
  const pt::Matrix4 r = pt::Rotate(pt::NewVector3(1, -2, 0.5), 1.234);
  pt::Matrix4 rt{};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) rt[4 * i + j] = r[4 * j + i];
  const pt::Matrix4 id = pt::MultiplyMatrix(r, rt);
  for (int k = 0; k < 16; ++k) EXPECT(std::fabs(id[k] - pt::IdentityMatrix()[k]) < 1e-15);
  EXPECT(pt::DegreesToRadians(180) == M_PI);
  EXPECT(std::fabs(pt::DefaultThreshold(64, 0.5, 2, 2) - 8) < 1e-15);  // 64 / 2^(1 + 2)
  // calcNormal: counter-clockwise corners seen from +z give +z at every vertex
  const pt::Primitive tri = pt::NewTriangleWithoutNormals({0, 0, 0}, {2, 0, 0}, {0, 3, 0});
  for (const auto& v : tri.Vertices) EXPECT(v.Normal.X == 0 && v.Normal.Y == 0 && v.Normal.Z == 1);
}

void case_camera() {
  // NewDefaultCamera looks from the origin along +x with +y up (camera.go:45-51)
  const double ar = 1.5, fov = 90;
  auto cam = pt::NewDefaultCamera(ar, fov);
  const trace_camera c = cam->Raw();
  const double h = 2 * std::tan(M_PI / 4);
  EXPECT(cam->W().X == -1 && cam->W().Y == 0 && cam->W().Z == 0);
  EXPECT(std::fabs(c.horizontal[2] - ar * h) < 1e-15 && c.horizontal[0] == 0 && c.horizontal[1] == 0);
  EXPECT(std::fabs(c.vertical[1] - h) < 1e-15 && c.vertical[0] == 0 && c.vertical[2] == 0);
  // the centre sample (0.5, 0.5) looks straight ahead
  for (int a = 0; a < 3; ++a) {
    const double d = c.lower_left[a] + 0.5 * c.horizontal[a] + 0.5 * c.vertical[a] - c.origin[a];
    EXPECT(std::fabs(d - (a == 0 ? 1.0 : 0.0)) < 1e-15);
  }
  // SetTransformation: looking from (0, 0, 5) at the origin, the centre ray points along -z
  cam->SetTransformation(pt::CameraTransformation{{0, 0, 5}, {0, 0, 0}, {0, 1, 0}});
  const trace_camera c2 = cam->Raw();
  EXPECT(c2.origin[2] == 5);
  for (int a = 0; a < 3; ++a) {
    const double d = c2.lower_left[a] + 0.5 * c2.horizontal[a] + 0.5 * c2.vertical[a] - c2.origin[a];
    EXPECT(std::fabs(d - (a == 2 ? -1.0 : 0.0)) < 1e-15);
  }
  // Translate moves the origin and the corner together
  cam->Translate(pt::NewVector3(1, 2, 3));
  const trace_camera c3 = cam->Raw();
  for (int a = 0; a < 3; ++a) EXPECT(std::fabs((c3.lower_left[a] - c2.lower_left[a]) - (a + 1)) < 1e-12);
}

void case_buffers() {
  // PixelBuffer: progressive sums fold into the running mean of every sample
  auto pb = pt::NewPxlBuffer(2, 1);
  Rng r{5};
  std::vector<double> samples;
  double incremental = 0;  // c += (s - c) / n, as Pixel.addSample does
  for (int pass = 0; pass < 3; ++pass) {
    const int spp = 1 + pass;
    float sum[6] = {0}, last[6] = {0};
    for (int s = 0; s < spp; ++s) {
      const float v = float(r.uniform());
      samples.push_back(v);
      incremental += (double(v) - incremental) / double(samples.size());
      sum[0] += v;
      sum[3] += 2 * v;
      last[0] = v;
    }
    pb->AddSamples(sum, last, spp);
  }
  EXPECT(pb->At(0, 0).Samples == 6 && pb->At(1, 0).Samples == 6);
  EXPECT_MSG(std::fabs(pb->At(0, 0).Mean.X - incremental) < 1e-7, "%.12g vs %.12g", pb->At(0, 0).Mean.X,
             incremental);
  EXPECT(std::fabs(pb->At(1, 0).Mean.X - 2 * incremental) < 2e-7);
  // FrameBuffer keeps the last pass's sample
  auto fb = pt::NewFrameBuffer(1, 1);
  float s1[3] = {5, 5, 5}, l1[3] = {1, 2, 3};
  fb->AddSamples(s1, l1, 4);
  EXPECT(fb->At(0, 0).X == 1 && fb->At(0, 0).Y == 2 && fb->At(0, 0).Z == 3);
  // the aspect-ratio constructors truncate the width (buffer.go:34-37, 77-80)
  EXPECT(pt::NewFrameBufferAR(64, 1.5)->Width() == 96);
  EXPECT(pt::NewPxlBufferAR(10, 1.99)->Width() == 19);
  expect_error(TRACE_E_INVALID, [] { pt::NewPxlBuffer(-1, 2); }, "negative width");
}

std::string temp_path(const char* name) {
  const char* d = std::getenv("TMPDIR");
  return std::string(d && *d ? d : "/tmp") + "/pt_test_" + std::to_string(long(getpid())) + "_" + name;
}

void case_obj(const std::string& golden) {
  // the reference's only fixture: a unit cube of six quads with per-face normals
  const pt::Geometry cube = pt::ParseFromPath(golden + "/cube.obj");
  EXPECT_MSG(cube.size() == 12, "%zu triangles", cube.size());
  for (const auto& t : cube) {
    pt::Vector3 centroid{};
    for (const auto& v : t.Vertices) {
      for (double x : {v.Position.X, v.Position.Y, v.Position.Z}) EXPECT(x == 0.5 || x == -0.5);
      centroid = pt::Add(centroid, pt::Mul(v.Position, 1.0 / 3));
    }
    // every corner carries the face normal: an axis, pointing out of the cube
    const pt::Vector3 n = t.Vertices[0].Normal;
    EXPECT(std::fabs(n.X) + std::fabs(n.Y) + std::fabs(n.Z) == 1);
    for (const auto& v : t.Vertices) EXPECT(v.Normal.X == n.X && v.Normal.Y == n.Y && v.Normal.Z == n.Z);
    EXPECT(pt::Dot(n, centroid) > 0);
    // and the geometric normal agrees with it
    const pt::Vector3 g =
        pt::Cross(pt::Sub(t.Vertices[1].Position, t.Vertices[0].Position), pt::Sub(t.Vertices[2].Position, t.Vertices[0].Position));
    EXPECT(pt::Dot(g, n) > 0);
  }
  // fans, negative indices, v//vn corners and skipped records
  const std::string p = temp_path("fan.obj");
  {
    std::ofstream f(p);
    f << "# comment\no thing\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0.5 1.5 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\n"
         "f 1 2 3 4 5\nf -3 -2 -1\nf 1//1 2//1 3//1\ns off\n";
  }
  const pt::Geometry g = pt::ParseFromPath(p);
  std::remove(p.c_str());
  EXPECT_MSG(g.size() == 5, "%zu triangles", g.size());
  // the pentagon as a fan from its first corner: (1 2 3), (1 3 4), (1 4 5)
  const double want_x[3][3] = {{0, 1, 1}, {0, 1, 0.5}, {0, 0.5, 0}};
  for (int k = 0; k < 3; ++k)
    for (int c = 0; c < 3; ++c) EXPECT(g[k].Vertices[c].Position.X == want_x[k][c]);
  // -3 -2 -1 are the last three vertices: 3, 4, 5
  EXPECT(g[3].Vertices[0].Position.X == 1 && g[3].Vertices[1].Position.X == 0.5 && g[3].Vertices[2].Position.X == 0);
  EXPECT(g[4].Vertices[1].Normal.Z == 1);
  // errors where the reference panics
  expect_error(TRACE_E_INVALID, [] { pt::ParseFromPath("/nonexistent/cube.obj"); }, "missing file");
  const std::string q = temp_path("bad.obj");
  {
    std::ofstream f(q);
    f << "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n";
  }
  expect_error(TRACE_E_INVALID, [&] { pt::ParseFromPath(q); }, "index out of range");
  {
    std::ofstream f(q);
    f << "v 0 0 zero\n";
  }
  expect_error(TRACE_E_INVALID, [&] { pt::ParseFromPath(q); }, "bad number");
  std::remove(q.c_str());
}

void case_no_device() {
  trace_ctx* c = nullptr;
  if (trace_ctx_create(0, &c) == TRACE_OK) {
    trace_ctx_destroy(c);
    std::printf("    (a device is present: the no-device error is not reachable)\n");
    return;
  }
  const pt::Tracables t = random_soup(10, 1);
  expect_error(TRACE_E_NODEVICE, [&] { pt::DefaultLBVH(t); }, "DefaultLBVH without a device");
  // SetDevice stays possible while no context exists
  pt::SetDevice(0);
}

// ---- device cases ------------------------------------------------------------------------

void case_lbvh() {
  const pt::Tracables prims = random_soup(30000, 11);
  const pt::BVH bvh = pt::DefaultLBVH(prims);
  const std::vector<double> pos = positions(prims);
  oracle_tree* ot = oracle_lbvh_build(pos.data(), int64_t(prims.size()), nullptr, TRACE_KEY_REF63);
  const int64_t on = oracle_tree_node_count(ot);
  std::vector<int32_t> orec(4 * on);
  std::vector<double> obox(6 * on);
  oracle_tree_canonical(ot, orec.data(), obox.data());
  int64_t n = 0;
  trace_ctx* c = pt::DeviceContext();
  EXPECT(trace_bvh_export_canonical(c, bvh.Handle(), 0, nullptr, nullptr, &n) == TRACE_OK);
  EXPECT_MSG(n == on, "%lld nodes vs %lld", (long long)n, (long long)on);
  std::vector<int32_t> rec(4 * n);
  std::vector<float> box(6 * n);
  EXPECT(trace_bvh_export_canonical(c, bvh.Handle(), n, rec.data(), box.data(), &n) == TRACE_OK);
  // kind and the covered range of every node in preorder (the topology), and its box
  for (int64_t k = 0; k < n; ++k)
    EXPECT_MSG(rec[4 * k] == orec[4 * k] && rec[4 * k + 1] == orec[4 * k + 1] && rec[4 * k + 2] == orec[4 * k + 2],
               "node %lld", (long long)k);
  for (int64_t k = 0; k < 6 * n; ++k) EXPECT(double(box[k]) == obox[k]);
  EXPECT(std::fabs(bvh.Cost() - oracle_tree_cost(ot)) <= 1e-12 * oracle_tree_cost(ot));
  EXPECT(bvh.Size() == on);
  // LBVH with the enclosing box given explicitly is the same tree
  const pt::BVH b2 = pt::LBVH(prims, pt::Enclosing(prims), 8);
  EXPECT(b2.Cost() == bvh.Cost());
  // closest hit against the oracle
  const std::vector<pt::Ray> rays = incoherent_rays(20000, 12);
  const std::vector<float> rf = ray_floats(rays, 0.001f, std::numeric_limits<float>::infinity());
  std::vector<int32_t> ohit(rays.size());
  std::vector<double> otv(rays.size());
  oracle_closest_hit(ot, rf.data(), int64_t(rays.size()), ohit.data(), otv.data(), kThreads);
  check_hits(bvh.ClosestHit(rays), ohit, otv);
  // traversalSteps (bvh.go:21-44) of the first 4000 rays, counted as the reference
  // counts them, with the reference's heat-map bounds (0.001, MaxFloat64)
  const std::vector<pt::Ray> sub(rays.begin(), rays.begin() + 4000);
  const double tmax = std::numeric_limits<double>::max();
  std::vector<double> r64;
  for (const auto& q : sub)
    r64.insert(r64.end(), {q.Origin.X, q.Origin.Y, q.Origin.Z, 0.001, q.Direction.X, q.Direction.Y, q.Direction.Z,
                           tmax});
  std::vector<int32_t> osteps(sub.size());
  oracle_traversal_steps(ot, r64.data(), int64_t(sub.size()), osteps.data(), kThreads);
  const std::vector<int32_t> steps = bvh.TraversalSteps(sub, 0.001, tmax);
  EXPECT(steps == osteps);
  EXPECT(*std::max_element(steps.begin(), steps.end()) > 10);
  oracle_tree_free(ot);
}

void case_phr() {
  const pt::Tracables prims = random_soup(30000, 21);
  const std::vector<double> pos = positions(prims);
  oracle_tree* ot = oracle_lbvh_build(pos.data(), int64_t(prims.size()), nullptr, TRACE_KEY_REF63);
  const pt::BVH aux = pt::DefaultLBVH(prims);
  const std::vector<pt::Ray> rays = incoherent_rays(20000, 22);
  const std::vector<float> rf = ray_floats(rays, 0.001f, std::numeric_limits<float>::infinity());
  for (int branching : {2, 4}) {
    std::printf("    branching %d\n", branching);
    std::fflush(stdout);
    // benchmark_test.go:21-22 builds NewPHRBuilder(prims, 0.5, 6, 4, threads).Build()
    pt::PhrBuilder builder = pt::NewPHRBuilder(prims, 0.5, 6, branching, 8);
    const pt::BVH bvh = builder.Build();
    oracle_phr* op = oracle_phr_build(ot, 0.5, 6, branching, 2500);
    const double oc = oracle_phr_cost(op);
    EXPECT_MSG(std::fabs(bvh.Cost() - oc) <= 1e-9 * oc, "Cost %.12g vs %.12g", bvh.Cost(), oc);
    // BuildFromAuxilary over an LBVH of the same tracables is the same tree
    const pt::BVH again = builder.BuildFromAuxilary(aux);
    EXPECT(again.Cost() == bvh.Cost() && again.Size() == bvh.Size());
    const auto with_cost = builder.BuildWithCost(aux);
    EXPECT_MSG(with_cost.second == oracle_phr_split_work(op), "build cost %lld vs %lld",
               (long long)with_cost.second, (long long)oracle_phr_split_work(op));
    std::vector<int32_t> ohit(rays.size());
    std::vector<double> otv(rays.size());
    oracle_phr_closest_hit(op, rf.data(), int64_t(rays.size()), ohit.data(), otv.data(), kThreads);
    check_hits(bvh.ClosestHit(rays), ohit, otv);
    oracle_phr_free(op);
  }
  // NewDefaultBuilder is alpha 0.5, delta 6, branching 2
  const pt::PhrBuilder d = pt::NewDefaultBuilder(prims);
  EXPECT(d.Alpha == 0.5 && d.Delta == 6 && d.BranchingFactor == 2 && d.MaxCutSize == 2500);
  // the device builds branching 2 to 8
  expect_error(TRACE_E_UNSUPPORTED, [&] { pt::NewPHRBuilder(prims, 0.5, 6, 1, 1).BuildFromAuxilary(aux); },
               "branching 1");
  expect_error(TRACE_E_INVALID, [] { pt::PhrBuilder().BuildFromAuxilary(pt::BVH()); }, "empty auxiliary tree");
  oracle_tree_free(ot);
}

void case_optimizers() {
  const pt::Tracables prims = random_soup(20000, 31);
  const pt::BVH aux = pt::DefaultLBVH(prims);
  // benchmark_test.go:53-67: NewDefaultGridOptimizer(res, res).OptimizedPHRparams(aux, branching, threads)
  const pt::GridOptimizer grid = pt::NewDefaultGridOptimizer(256, 256);
  const auto ad = grid.OptimizedPHRparams(aux, 2, 8);
  bool on_grid = false;
  for (double a : grid.Alphas)
    for (double d : grid.Deltas) on_grid = on_grid || (a == ad.first && d == ad.second);
  EXPECT_MSG(on_grid, "(%g, %g) not on the grid", ad.first, ad.second);
  // the same call through the C-ABI gives the same pair
  double a = 0, d = 0;
  trace_ctx* c = pt::DeviceContext();
  EXPECT(trace_phr_optimize_grid(c, aux.Handle(), grid.Alphas.data(), 4, grid.Deltas.data(), 4, 2, 2500, 256, 256,
                                 nullptr, nullptr, &a, &d) == TRACE_OK);
  EXPECT(a == ad.first && d == ad.second);
  pt::BayesianOptimizer bo = pt::NewDefaultBayesianOptimizer(256, 256);
  bo.Rounds = 8;
  bo.RandomRounds = 3;
  const auto bd = bo.OptimizedPHRparams(aux, 2, 8);
  EXPECT(bd.first >= 0.4 && bd.first <= 0.55 && bd.second >= 6 && bd.second <= 9);
  const auto bd2 = bo.OptimizedPHRparams(aux, 2, 8);  // the same seed, the same search
  EXPECT(bd2 == bd);
}

// A small room: the reference's cube (rotated, lifted, Reflective or Diffuse), a
// floor, a light above, as a scene graph with a nested node.
struct Room {
  std::shared_ptr<pt::Scene> scene;
  // the graph as trace_scene_compile / oracle_scene_flatten take it, written out by hand
  std::vector<int32_t> parent, node_mesh;
  std::vector<double> matrix;
  std::vector<double> mpos, mnrm;
  std::vector<int64_t> off;
  std::vector<float> mats8;
};

pt::Geometry quad(double y, double s, bool down) {
  const pt::Vector3 a{-s, y, -s}, b{s, y, -s}, c{s, y, s}, d{-s, y, s};
  if (down) return {pt::NewTriangleWithoutNormals(a, b, c), pt::NewTriangleWithoutNormals(a, c, d)};
  return {pt::NewTriangleWithoutNormals(a, c, b), pt::NewTriangleWithoutNormals(a, d, c)};
}

// exact: no rotation and power-of-two steps only, so every flattened corner is an
// FP32 value (the device's FP32 node boxes then equal the FP64 ones, and walks
// that graze a box count the same steps).
Room make_room(const std::string& golden, pt::Material cube_mat, bool exact = false) {
  Room r;
  const pt::Geometry cube = pt::ParseFromPath(golden + "/cube.obj");
  auto cube_mesh = pt::NewMesh(cube, cube_mat);
  auto floor_mesh = pt::NewMesh(quad(0, 3, false), pt::Diffuse{pt::NewColor(0.8, 0.8, 0.8)});
  auto light_mesh = pt::NewMesh(quad(2.5, exact ? 0.75 : 0.8, true), pt::Light{pt::NewColor(6, 6, 6)});
  r.scene = pt::NewScene();
  auto floor = pt::NewSceneNode(floor_mesh);
  auto group = pt::NewSceneNode(nullptr);
  group->Translate(0, 0.5, 0);
  auto box = pt::NewSceneNode(cube_mesh);
  const double turn = exact ? 0.0 : 0.4, s1 = exact ? 0.75 : 0.9;
  box->Rotate(pt::NewVector3(0, 1, 0), turn);
  box->ScaleUniform(s1);
  auto box2 = pt::NewSceneNode(cube_mesh);  // the same mesh placed twice
  const double tx = exact ? 1.5 : 1.4, ty = exact ? -0.25 : -0.2, tz = exact ? 0.25 : 0.3, sy = exact ? 0.5 : 0.6;
  box2->Translate(tx, ty, tz);
  box2->Scale(0.5, sy, 0.5);
  auto light = pt::NewSceneNode(light_mesh);
  group->Add(box);
  group->Add(box2);
  r.scene->Add(floor);
  r.scene->Add(group);
  r.scene->Add(light);
  // the same graph by hand: nodes floor, group, box, box2, light; meshes floor 0,
  // cube 1, light 2 (first-use order)
  r.parent = {-1, -1, 1, 1, -1};
  r.node_mesh = {0, -1, 1, 1, 2};
  const pt::Matrix4 id = pt::IdentityMatrix();
  const pt::Matrix4 mg = pt::MultiplyMatrix(id, pt::Translate(0, 0.5, 0));
  const pt::Matrix4 mb = pt::MultiplyMatrix(pt::MultiplyMatrix(id, pt::Rotate(pt::NewVector3(0, 1, 0), turn)),
                                            pt::ScaleUniform(s1));
  const pt::Matrix4 mb2 = pt::MultiplyMatrix(pt::MultiplyMatrix(id, pt::Translate(tx, ty, tz)),
                                             pt::Scale(0.5, sy, 0.5));
  for (const pt::Matrix4* m : {&id, &mg, &mb, &mb2, &id}) r.matrix.insert(r.matrix.end(), m->begin(), m->end());
  r.off = {0};
  for (const auto* mesh : {floor_mesh.get(), cube_mesh.get(), light_mesh.get()}) {
    for (const auto& p : mesh->geometry)
      for (const auto& v : p.Vertices) {
        r.mpos.insert(r.mpos.end(), {v.Position.X, v.Position.Y, v.Position.Z});
        r.mnrm.insert(r.mnrm.end(), {v.Normal.X, v.Normal.Y, v.Normal.Z});
      }
    r.off.push_back(r.off.back() + int64_t(mesh->geometry.size()));
    const pt::Material& m = mesh->material;
    r.mats8.insert(r.mats8.end(), {float(m.Kind), float(m.Albedo.X), float(m.Albedo.Y), float(m.Albedo.Z),
                                   float(m.Param), 0, 0, 0});
  }
  return r;
}

struct OracleScene {
  std::vector<double> pos, nrm;
  std::vector<int32_t> mat;
  oracle_tree* tree = nullptr;
  ~OracleScene() {
    if (tree) oracle_tree_free(tree);
  }
};

void flatten_oracle(const Room& r, OracleScene& o) {
  const int32_t mesh_mat[3] = {0, 1, 2};
  const int64_t n = oracle_scene_flatten(r.mpos.data(), r.mnrm.data(), r.off.data(), mesh_mat, r.parent.data(),
                                         r.node_mesh.data(), r.matrix.data(), int32_t(r.parent.size()), nullptr,
                                         nullptr, nullptr);
  o.pos.resize(9 * n);
  o.nrm.resize(9 * n);
  o.mat.resize(n);
  oracle_scene_flatten(r.mpos.data(), r.mnrm.data(), r.off.data(), mesh_mat, r.parent.data(), r.node_mesh.data(),
                       r.matrix.data(), int32_t(r.parent.size()), o.pos.data(), o.nrm.data(), o.mat.data());
  o.tree = oracle_lbvh_build(o.pos.data(), n, nullptr, TRACE_KEY_REF63);
}

// Fraction of pixels whose means differ by more than 3 sigma (the oracle's sample
// variance for both) plus a small floor.
double frac_beyond_3_sigma(const std::vector<double>& m_gpu, const std::vector<double>& s_or,
                           const std::vector<double>& q_or, int n) {
  int64_t bad = 0;
  for (size_t i = 0; i < m_gpu.size(); ++i) {
    const double m = s_or[i] / n, v = std::max(q_or[i] / n - m * m, 0.0) * n / std::max(n - 1, 1);
    const double se = std::sqrt(2 * v / n);
    bad += std::fabs(m_gpu[i] - m) > 3 * se + 1e-4 + 1e-5 * std::fabs(m);
  }
  return double(bad) / double(m_gpu.size());
}

void case_scene_and_render(const std::string& golden) {
  const char* names[3] = {"Diffuse", "Reflective", "Refractive"};
  for (int specular = 0; specular < 3; ++specular) {
    std::printf("    cube material %s\n", names[specular]);
    std::fflush(stdout);
    const pt::Material cube_mat = specular == 2   ? pt::Material(pt::Refractive{pt::NewColor(0.95, 0.95, 1.0), 1.5})
                                  : specular == 1 ? pt::Material(pt::Reflective{pt::NewColor(0.9, 0.8, 0.6), 0.2})
                                                  : pt::Material(pt::Diffuse{pt::NewColor(0.7, 0.3, 0.3)});
    // the sun miss shader with the glass cube, the sky otherwise
    const pt::MissShader miss = specular == 2 ? pt::SunMissShader : pt::SkyMissShader;
    Room room = make_room(golden, cube_mat);
    OracleScene o;
    flatten_oracle(room, o);
    const int64_t n = int64_t(o.mat.size());
    EXPECT(int64_t(room.scene->UntransformedTracables().size()) == n);
    const pt::BVH bvh = room.scene->CompileLBVH();
    // the device-flattened triangles are the oracle's, bit for bit
    std::vector<double> dpos(9 * n);
    EXPECT(trace_mesh_export(pt::DeviceContext(), bvh.MeshHandle(), dpos.data()) == TRACE_OK);
    EXPECT(dpos == o.pos);
    std::vector<float> dn(9 * n);
    std::vector<int32_t> dm(n);
    EXPECT(trace_mesh_export_shading(pt::DeviceContext(), bvh.MeshHandle(), dn.data(), dm.data()) == TRACE_OK);
    EXPECT(dm == o.mat);
    for (int64_t k = 0; k < 9 * n; ++k) EXPECT(dn[k] == float(o.nrm[k]));

    // benchmark_test.go:120-137: a FrameBuffer, the default camera moved to a view,
    // the benchmark renderer (1 spp, white miss), one RenderToBuffer
    const double AR = 1.0, FOV = 60.0;
    auto buff = pt::NewFrameBufferAR(48, AR);
    auto camera = pt::NewDefaultCamera(AR, FOV);
    camera->SetTransformation(pt::CameraTransformation{{3.2, 1.8, 3.6}, {0, 0.6, 0}, {0, 1, 0}});
    auto renderer = pt::NewBenchmarkRenderer(bvh, camera);
    renderer->Seed = 7;
    renderer->RenderToBuffer(*buff);
    const int W = buff->Width(), H = buff->Height();
    const trace_camera cam = camera->Raw();
    const double* cp = cam.origin;  // origin, lower_left, horizontal, vertical: 12 doubles in order
    std::vector<double> s(3 * W * H), q(3 * W * H), last(3 * W * H);
    oracle_render_ex(o.tree, o.nrm.data(), o.mat.data(), room.mats8.data(), 3, cp, W, H, 1, 5,
                     TRACE_MISS_WHITE, 7, kThreads, s.data(), q.data(), last.data());
    int64_t same = 0;
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        const pt::Color g = buff->At(x, y);
        const double* e = &last[3 * (size_t(y) * W + x)];
        same += std::fabs(g.X - e[0]) <= 1e-3 * (1 + std::fabs(e[0])) &&
                std::fabs(g.Y - e[1]) <= 1e-3 * (1 + std::fabs(e[1])) &&
                std::fabs(g.Z - e[2]) <= 1e-3 * (1 + std::fabs(e[2]));
      }
    // the same Philox stream per (seed, path, bounce): paths differ only where FP32
    // geometry sends them elsewhere
    const double frac_same = double(same) / double(W * H);
    double mean = 0;
    for (double v : last) mean += v / double(last.size());
    std::printf("      1 spp: %.4f of pixels equal the oracle's (oracle mean %.4f)\n", frac_same, mean);
    EXPECT_MSG(frac_same >= 0.97, "%.4f of 1-spp pixels equal the oracle's", frac_same);
    EXPECT(mean > 0.05);  // the light and the white sky reach the camera

    // progressive rendering into a PixelBuffer: two calls of 8 samples hold the
    // running mean of samples 0..15, within 3 sigma of a 16-sample oracle render
    auto pb = pt::NewPxlBuffer(W, H);
    auto r2 = pt::NewDefaultRenderer(bvh, camera);
    EXPECT(r2->Spp == 300 && r2->MaxDepth == 5 && r2->Miss == pt::DefaultMissShader);
    r2->Spp = 8;
    r2->Miss = miss;
    r2->Seed = 9;
    r2->RenderToBuffer(*pb);
    r2->RenderToBuffer(*pb);
    EXPECT(r2->SamplesDrawn() == 16 && pb->At(0, 0).Samples == 16 && pb->At(W - 1, H - 1).Samples == 16);
    oracle_render_ex(o.tree, o.nrm.data(), o.mat.data(), room.mats8.data(), 3, cp, W, H, 16, 5, int32_t(miss), 9,
                     kThreads, s.data(), q.data(), last.data());
    std::vector<double> m(3 * W * H);
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        const pt::Color c = pb->At(x, y).Mean;
        double* d = &m[3 * (size_t(y) * W + x)];
        d[0] = c.X, d[1] = c.Y, d[2] = c.Z;
      }
    const double bad = frac_beyond_3_sigma(m, s, q, 16);
    std::printf("      16 spp in two calls: %.4f of pixel channels beyond 3 sigma\n", bad);
    EXPECT_MSG(bad <= 0.01, "%.4f of pixel channels beyond 3 sigma", bad);

    // Compile / CompilePHR give trees over the same flattened triangles
    const pt::BVH phr = room.scene->CompilePHR(0.55, 9, 4);
    const pt::BVH def = room.scene->Compile();
    EXPECT(phr.Info().n_prims == n && def.Info().n_prims == n && def.Info().arity == 2);
    auto fb2 = pt::NewFrameBufferAR(48, AR);
    auto r3 = pt::NewRealtimeRenderer(phr, camera);
    EXPECT(r3->Spp == 1 && r3->Miss == pt::SkyMissShader);
    r3->Seed = 7;
    r3->Miss = pt::WhiteMissShader;
    r3->RenderToBuffer(*fb2);
    // the PHR tree finds the same closest hits, so its 1-spp frame matches the LBVH's
    int64_t same2 = 0;
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        const pt::Color a = fb2->At(x, y), b = buff->At(x, y);
        same2 += a.X == b.X && a.Y == b.Y && a.Z == b.Z;
      }
    std::printf("      PHR (0.55, 9, 4) frame: %.4f of pixels equal the LBVH frame\n", double(same2) / double(W * H));
    EXPECT_MSG(double(same2) / double(W * H) >= 0.99, "%.4f of pixels equal", double(same2) / double(W * H));
  }
  expect_error(TRACE_E_INVALID, [] { pt::NewScene()->CompileLBVH(); }, "empty scene");
}

void case_heatmap(const std::string& golden) {
  Room room = make_room(golden, pt::Diffuse{pt::NewColor(0.5, 0.5, 0.5)}, true);
  OracleScene o;
  flatten_oracle(room, o);
  const pt::BVH bvh = room.scene->CompileLBVH();
  auto camera = pt::NewCamera(1.25, 50, pt::CameraTransformation{{3, 2, 4}, {0, 0.5, 0}, {0, 1, 0}});
  auto pb = pt::NewPxlBufferAR(40, 1.25);
  auto hm = pt::NewHeatMapRenderer(bvh, camera, 6);
  hm->RenderToBuffer(*pb);
  const int W = pb->Width(), H = pb->Height();
  const trace_camera cam = camera->Raw();
  std::vector<float> rgb(3 * W * H);
  oracle_heatmap(o.tree, cam.origin, W, H, 6, rgb.data(), nullptr, kThreads);
  int64_t same = 0;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const pt::Color c = pb->At(x, y).Mean;
      const float* e = &rgb[3 * (size_t(y) * W + x)];
      same += float(c.X) == e[0] && float(c.Y) == e[1] && float(c.Z) == e[2];
    }
  // every colour equal: the steps are counted by the reference's own walk in FP64
  EXPECT_MSG(same == int64_t(W) * H, "%.4f of heat-map pixels equal", double(same) / double(W * H));
  // DefaultTraversalCountShader's two branches (render.go:250-258)
  const pt::Color over = pt::DefaultTraversalCountShader(10, 15), under = pt::DefaultTraversalCountShader(10, 5);
  EXPECT(over.X == 3 && over.Y == 0 && over.Z == 0);
  EXPECT(under.X == 0 && under.Y == 0.4 && under.Z == 0.6);
}

void case_errors() {
  // an empty tree, an empty tracable list, a tracable list without materials
  auto cam = pt::NewDefaultCamera(1, 60);
  auto fb = pt::NewFrameBuffer(8, 8);
  expect_error(TRACE_E_INVALID, [&] { pt::NewBenchmarkRenderer(pt::BVH(), cam)->RenderToBuffer(*fb); },
               "render without a tree");
  expect_error(TRACE_E_INVALID, [] { pt::DefaultLBVH(pt::Tracables{}); }, "no primitives");
  pt::Tracables t = random_soup(10, 3);
  t.MaterialIndex.pop_back();
  expect_error(TRACE_E_INVALID, [&] { pt::DefaultLBVH(t); }, "missing material index");
  t = random_soup(10, 3);
  t.MaterialIndex[0] = 5;
  expect_error(TRACE_E_INVALID, [&] { pt::DefaultLBVH(t); }, "material index out of range");
  expect_error(TRACE_E_INVALID, [] { pt::SetDevice(1); }, "SetDevice after first use");
  // spheres and triangles together: the sphere's hits come back with its index
  pt::Tracables s = random_soup(200, 4);
  s.Prims.push_back(pt::NewSphere(pt::NewVector3(5, 5, 5), 0.5));
  s.MaterialIndex.push_back(0);
  const pt::BVH b = pt::DefaultLBVH(s);
  const auto h = b.ClosestHit({pt::Ray{{5, 5, 0}, {0, 0, 1}}});
  EXPECT(h[0].Ok && h[0].Index == 200 && std::fabs(h[0].T - 4.5) < 1e-6);
  const auto miss = b.ClosestHit({pt::Ray{{5, 5, 0}, {0, 0, -1}}});
  EXPECT(!miss[0].Ok && miss[0].Index == -1);
}

struct Case {
  const char* name;
  bool gpu;
  std::function<void()> fn;
};

}  // namespace

int main(int argc, char** argv) {
  bool gpu = false;
  std::string filter;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--gpu") gpu = true;
    else if (a == "--cpu") gpu = false;
    else filter = a;
  }
  const char* g = std::getenv("PT_TEST_GOLDEN");
  const std::string golden = g && *g ? g : "tests/golden";
  const std::vector<Case> cases = {
      {"math", false, case_math},
      {"camera", false, case_camera},
      {"buffers", false, case_buffers},
      {"obj", false, [&] { case_obj(golden); }},
      {"no_device", false, case_no_device},
      {"lbvh", true, case_lbvh},
      {"phr", true, case_phr},
      {"optimizers", true, case_optimizers},
      {"scene_and_render", true, [&] { case_scene_and_render(golden); }},
      {"heatmap", true, [&] { case_heatmap(golden); }},
      {"errors", true, case_errors},
  };
  int failed = 0, ran = 0;
  for (const Case& c : cases) {
    if (c.gpu && !gpu) continue;
    if (gpu && std::string(c.name) == "no_device") continue;
    if (!filter.empty() && std::string(c.name).find(filter) == std::string::npos) continue;
    std::printf("RUN  %s\n", c.name);
    std::fflush(stdout);
    const auto t0 = std::chrono::steady_clock::now();
    try {
      c.fn();
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      std::printf("PASS %s (%.0f ms)\n", c.name, ms);
    } catch (const Failure& f) {
      ++failed;
      std::printf("FAIL %s: %s\n", c.name, f.msg.c_str());
    } catch (const std::exception& e) {
      ++failed;
      std::printf("FAIL %s: exception: %s\n", c.name, e.what());
    }
    std::fflush(stdout);
    ++ran;
  }
  std::printf("%d of %d cases passed\n", ran - failed, ran);
  return failed;
}
