// The pkg/pt host API (trace_pt.hpp) over the C-ABI of trace_hip.h.
//
// Host arithmetic (vectors, matrices, the camera, triangle normals) is FP64 with
// every operation rounded on its own and in the reference's operation order, so
// the values handed to the device equal the ones the reference computes (built
// with -ffp-contract=off).  Everything else is marshalling: tracables become the
// device's primitive slots (9 doubles each) plus a per-primitive material index
// and a material table, trees are reference-counted handles that keep the device
// mesh they were built from alive, and renders fold the device's per-pixel sums
// into the caller's buffer.
#include "trace_pt.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>
#include <unordered_map>

namespace pt {

namespace {

// One context per process; recursive so that a handle released while a call holds
// the lock (an exception unwinding through it) can still take it.  Never destroyed,
// so that handles held by the caller's own static objects can still be released
// after this file's statics would have been.
std::recursive_mutex& g_mu = *new std::recursive_mutex;
trace_ctx* g_ctx = nullptr;
int g_device = 0;

using Lock = std::lock_guard<std::recursive_mutex>;

// g_mu held
trace_ctx* ctx_locked() {
  if (!g_ctx) {
    trace_ctx* c = nullptr;
    const int rc = trace_ctx_create(g_device, &c);
    if (rc != TRACE_OK)
      throw Error(rc, "pt: no usable HIP device " + std::to_string(g_device) + " (trace_ctx_create returned " +
                          std::to_string(rc) + ")");
    g_ctx = c;
  }
  return g_ctx;
}

void check(trace_ctx* c, int rc, const char* what) {
  if (rc != TRACE_OK) throw Error(rc, std::string("pt: ") + what + ": " + trace_last_error(c));
}

[[noreturn]] void fail(int code, const std::string& msg) { throw Error(code, "pt: " + msg); }

}  // namespace

void SetDevice(int device) {
  Lock l(g_mu);
  if (g_ctx) fail(TRACE_E_INVALID, "SetDevice after the device context was created");
  g_device = device;
}

trace_ctx* DeviceContext() {
  Lock l(g_mu);
  return ctx_locked();
}

// ---- math ---------------------------------------------------------------------------

Vector3 Add(Vector3 a, Vector3 b) { return Vector3{a.X + b.X, a.Y + b.Y, a.Z + b.Z}; }
Vector3 Sub(Vector3 a, Vector3 b) { return Vector3{a.X - b.X, a.Y - b.Y, a.Z - b.Z}; }
Vector3 Mul(Vector3 a, double s) { return Vector3{a.X * s, a.Y * s, a.Z * s}; }
double Dot(Vector3 a, Vector3 b) { return a.X * b.X + a.Y * b.Y + a.Z * b.Z; }
Vector3 Cross(Vector3 a, Vector3 b) {
  return Vector3{a.Y * b.Z - a.Z * b.Y, a.Z * b.X - a.X * b.Z, a.X * b.Y - a.Y * b.X};
}
double Length(Vector3 a) { return std::sqrt(a.X * a.X + a.Y * a.Y + a.Z * a.Z); }
Vector3 Unit(Vector3 a) {
  const double l = Length(a);
  return Vector3{a.X / l, a.Y / l, a.Z / l};
}

Matrix4 IdentityMatrix() { return Matrix4{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}; }

Matrix4 MultiplyMatrix(const Matrix4& a, const Matrix4& b) {
  Matrix4 m{};
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c)
      m[4 * r + c] = a[4 * r] * b[c] + a[4 * r + 1] * b[4 + c] + a[4 * r + 2] * b[8 + c] + a[4 * r + 3] * b[12 + c];
  return m;
}

Matrix4 Scale(double x, double y, double z) { return Matrix4{x, 0, 0, 0, 0, y, 0, 0, 0, 0, z, 0, 0, 0, 0, 1}; }
Matrix4 ScaleUniform(double factor) { return Scale(factor, factor, factor); }
Matrix4 Translate(double x, double y, double z) { return Matrix4{1, 0, 0, x, 0, 1, 0, y, 0, 0, 1, z, 0, 0, 0, 1}; }

Matrix4 Rotate(Vector3 dir, double angle) {
  const Vector3 u = Unit(dir);
  const double w = std::cos(angle / 2), s = std::sin(angle / 2);
  const double x = s * u.X, y = s * u.Y, z = s * u.Z;
  // the rotation matrix of the unit quaternion (w, x, y, z)
  return Matrix4{1 - 2 * y * y - 2 * z * z, 2 * x * y - 2 * z * w,     2 * x * z + 2 * y * w,     0,
                 2 * x * y + 2 * z * w,     1 - 2 * x * x - 2 * z * z, 2 * y * z - 2 * x * w,     0,
                 2 * x * z - 2 * y * w,     2 * y * z + 2 * x * w,     1 - 2 * x * x - 2 * y * y, 0,
                 0,                         0,                         0,                         1};
}

double DegreesToRadians(double degree) { return degree * (M_PI / 180); }

// ---- primitives ---------------------------------------------------------------------------

Primitive NewTriangle(const std::array<Vertex, 3>& vertices) {
  Primitive p;
  p.Vertices = vertices;
  return p;
}

namespace {
// The unit normal at `point` of the triangle (point, right, left).
Vector3 calc_normal(Vector3 point, Vector3 right, Vector3 left) {
  const Vector3 pa = Sub(left, point), pb = Sub(right, point);
  return Unit(Cross(pb, pa));
}
}  // namespace

Primitive NewTriangleWithoutNormals(Vector3 v0, Vector3 v1, Vector3 v2) {
  return NewTriangle({Vertex{v0, calc_normal(v0, v1, v2)}, Vertex{v1, calc_normal(v1, v2, v0)},
                      Vertex{v2, calc_normal(v2, v0, v1)}});
}

Primitive NewSphere(Vector3 center, double radius) {
  Primitive p;
  p.IsSphere = true;
  p.Center = center;
  p.Radius = radius;
  return p;
}

namespace {

std::vector<std::string> split(const std::string& s, char sep) {
  std::vector<std::string> out;
  size_t b = 0;
  for (;;) {
    const size_t e = s.find(sep, b);
    out.push_back(s.substr(b, e == std::string::npos ? std::string::npos : e - b));
    if (e == std::string::npos) return out;
    b = e + 1;
  }
}

// strconv.Atoi: an optional sign and decimal digits, nothing else.
long long parse_int(const std::string& s, const std::string& path, long line) {
  size_t i = (!s.empty() && (s[0] == '+' || s[0] == '-')) ? 1 : 0;
  bool ok = i < s.size();
  for (size_t k = i; k < s.size(); ++k) ok = ok && s[k] >= '0' && s[k] <= '9';
  errno = 0;
  const long long v = ok ? std::strtoll(s.c_str(), nullptr, 10) : 0;
  if (!ok || errno == ERANGE)
    fail(TRACE_E_INVALID, path + ":" + std::to_string(line) + ": invalid index \"" + s + "\"");
  return v;
}

double parse_float(const std::string& s, const std::string& path, long line) {
  char* end = nullptr;
  const double v = std::strtod(s.c_str(), &end);
  if (s.empty() || end != s.c_str() + s.size())
    fail(TRACE_E_INVALID, path + ":" + std::to_string(line) + ": invalid number \"" + s + "\"");
  return v;
}

}  // namespace

Geometry ParseFromPath(const std::string& path) {
  std::ifstream in(path);
  if (!in) fail(TRACE_E_INVALID, "open " + path + ": no such file or not readable");
  // index 0 is a placeholder: OBJ indices start at 1, negative ones count from the end
  std::vector<Vector3> verts(1), normals(1);
  Geometry tris;
  std::string text;
  long line = 0;
  while (std::getline(in, text)) {
    ++line;
    std::istringstream fs(text);
    std::vector<std::string> f;
    for (std::string w; fs >> w;) f.push_back(w);
    if (f.empty()) continue;
    const std::string& key = f[0];
    if (key == "v" || key == "vn") {
      if (f.size() < 4) fail(TRACE_E_INVALID, path + ":" + std::to_string(line) + ": fewer than 3 coordinates");
      double x[3];
      for (int k = 0; k < 3; ++k) x[k] = parse_float(f[1 + k], path, line);
      for (size_t k = 4; k < f.size(); ++k) parse_float(f[k], path, line);  // extra values must still parse
      (key == "v" ? verts : normals).push_back(Vector3{x[0], x[1], x[2]});
    } else if (key == "f") {
      std::vector<long long> vi, ni;
      for (size_t a = 1; a < f.size(); ++a) {
        const std::vector<std::string> parts = split(f[a], '/');
        long long v = parse_int(parts[0], path, line);
        if (v < 0) v += static_cast<long long>(verts.size());
        vi.push_back(v);
        if (parts.size() >= 3) {
          long long n = parse_int(parts[2], path, line);
          if (n < 0) n += static_cast<long long>(normals.size());
          ni.push_back(n);
        }
      }
      const bool with_normals = ni.size() == vi.size();
      auto vert = [&](long long i) {
        if (i < 0 || i >= static_cast<long long>(verts.size()))
          fail(TRACE_E_INVALID, path + ":" + std::to_string(line) + ": vertex index out of range");
        return verts[size_t(i)];
      };
      auto nrm = [&](long long i) {
        if (i < 0 || i >= static_cast<long long>(normals.size()))
          fail(TRACE_E_INVALID, path + ":" + std::to_string(line) + ": normal index out of range");
        return normals[size_t(i)];
      };
      for (size_t i = 1; i + 2 <= vi.size(); ++i) {
        if (with_normals)
          tris.push_back(NewTriangle({Vertex{vert(vi[0]), nrm(ni[0])}, Vertex{vert(vi[i]), nrm(ni[i])},
                                      Vertex{vert(vi[i + 1]), nrm(ni[i + 1])}}));
        else
          tris.push_back(NewTriangleWithoutNormals(vert(vi[0]), vert(vi[i]), vert(vi[i + 1])));
      }
    }
  }
  return tris;
}

void Tracables::Append(const Tracables& other) {
  const int32_t base = static_cast<int32_t>(Materials.size());
  Prims.insert(Prims.end(), other.Prims.begin(), other.Prims.end());
  for (int32_t m : other.MaterialIndex) MaterialIndex.push_back(base + m);
  Materials.insert(Materials.end(), other.Materials.begin(), other.Materials.end());
}

AABB Enclosing(const Tracables& prims) {
  const double inf = std::numeric_limits<double>::infinity();
  AABB b{Vector3{inf, inf, inf}, Vector3{-inf, -inf, -inf}};
  for (const Primitive& p : prims.Prims) {
    Vector3 lo, hi;
    if (p.IsSphere) {
      const Vector3 r{p.Radius, p.Radius, p.Radius};
      lo = Sub(p.Center, r);
      hi = Add(p.Center, r);
    } else {
      const Vector3 &a = p.Vertices[0].Position, &c = p.Vertices[1].Position, &d = p.Vertices[2].Position;
      lo = Vector3{std::fmin(a.X, std::fmin(c.X, d.X)), std::fmin(a.Y, std::fmin(c.Y, d.Y)),
                   std::fmin(a.Z, std::fmin(c.Z, d.Z))};
      hi = Vector3{std::fmax(a.X, std::fmax(c.X, d.X)), std::fmax(a.Y, std::fmax(c.Y, d.Y)),
                   std::fmax(a.Z, std::fmax(c.Z, d.Z))};
    }
    b.Min = Vector3{std::fmin(b.Min.X, lo.X), std::fmin(b.Min.Y, lo.Y), std::fmin(b.Min.Z, lo.Z)};
    b.Max = Vector3{std::fmax(b.Max.X, hi.X), std::fmax(b.Max.Y, hi.Y), std::fmax(b.Max.Z, hi.Z)};
  }
  return b;
}

// ---- device meshes and trees -------------------------------------------------------------

namespace {

struct MeshState {
  trace_mesh* mesh = nullptr;
  int64_t n = 0;
  ~MeshState() {
    if (mesh) {
      Lock l(g_mu);
      trace_mesh_destroy(mesh);
    }
  }
};

// Primitive slots as the device takes them: a triangle's three corners, or a
// sphere's centre, radius and NaN (then four ignored values).
void pack(const Tracables& t, std::vector<double>& pos, std::vector<double>& nrm) {
  const size_t n = t.size();
  pos.assign(9 * n, 0.0);
  nrm.assign(9 * n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    const Primitive& p = t.Prims[i];
    double* q = &pos[9 * i];
    if (p.IsSphere) {
      q[0] = p.Center.X, q[1] = p.Center.Y, q[2] = p.Center.Z, q[3] = p.Radius;
      q[4] = std::numeric_limits<double>::quiet_NaN();
      continue;
    }
    for (int k = 0; k < 3; ++k) {
      const Vertex& v = p.Vertices[k];
      q[3 * k] = v.Position.X, q[3 * k + 1] = v.Position.Y, q[3 * k + 2] = v.Position.Z;
      nrm[9 * i + 3 * k] = v.Normal.X, nrm[9 * i + 3 * k + 1] = v.Normal.Y, nrm[9 * i + 3 * k + 2] = v.Normal.Z;
    }
  }
}

std::vector<float> material_table8(const std::vector<Material>& mats) {
  std::vector<float> out(8 * mats.size(), 0.0f);
  for (size_t m = 0; m < mats.size(); ++m) {
    out[8 * m] = float(mats[m].Kind);
    out[8 * m + 1] = float(mats[m].Albedo.X);
    out[8 * m + 2] = float(mats[m].Albedo.Y);
    out[8 * m + 3] = float(mats[m].Albedo.Z);
    out[8 * m + 4] = float(mats[m].Param);
  }
  return out;
}

std::shared_ptr<MeshState> upload(const Tracables& t) {
  if (t.MaterialIndex.size() != t.Prims.size())
    fail(TRACE_E_INVALID, "tracables: " + std::to_string(t.MaterialIndex.size()) + " material indices for " +
                              std::to_string(t.Prims.size()) + " primitives");
  std::vector<double> pos, nrm;
  pack(t, pos, nrm);
  const std::vector<float> mats = material_table8(t.Materials);
  Lock l(g_mu);
  trace_ctx* c = ctx_locked();
  auto ms = std::make_shared<MeshState>();
  check(c, trace_mesh_create(c, pos.data(), int64_t(t.size()), &ms->mesh), "mesh upload");
  ms->n = int64_t(t.size());
  check(c,
        trace_mesh_set_shading_ex(c, ms->mesh, nrm.data(), t.MaterialIndex.data(), mats.data(),
                                  int32_t(t.Materials.size())),
        "mesh shading");
  return ms;
}

}  // namespace

struct BVH::State {
  std::shared_ptr<MeshState> mesh;
  trace_bvh* bvh = nullptr;
  ~State() {
    if (bvh) {
      Lock l(g_mu);
      trace_bvh_destroy(bvh);
    }
  }
};

namespace {

const BVH::State& state_of(const std::shared_ptr<BVH::State>& s, const char* what) {
  if (!s) fail(TRACE_E_INVALID, std::string(what) + ": empty BVH");
  return *s;
}

}  // namespace

struct BvhAccess {
  static const std::shared_ptr<BVH::State>& of(const BVH& b) { return b.s_; }
};

namespace {

BVH lbvh_of(std::shared_ptr<MeshState> ms, const double* enclosing6) {
  auto st = std::make_shared<BVH::State>();
  st->mesh = std::move(ms);
  Lock l(g_mu);
  trace_ctx* c = ctx_locked();
  check(c, trace_lbvh_build(c, st->mesh->mesh, enclosing6, TRACE_KEY_REF63, &st->bvh), "LBVH build");
  return BVH(st);
}

}  // namespace

double BVH::Cost() const {
  const State& s = state_of(s_, "Cost");
  Lock l(g_mu);
  trace_ctx* c = ctx_locked();
  double v = 0;
  check(c, trace_bvh_cost(c, s.bvh, &v), "Cost");
  return v;
}

int64_t BVH::Size() const {
  const State& s = state_of(s_, "Size");
  Lock l(g_mu);
  trace_ctx* c = ctx_locked();
  int64_t v = 0;
  check(c, trace_bvh_size(c, s.bvh, &v), "Size");
  return v;
}

trace_bvh_info BVH::Info() const {
  const State& s = state_of(s_, "Info");
  Lock l(g_mu);
  trace_ctx* c = ctx_locked();
  trace_bvh_info info{};
  check(c, trace_bvh_get_info(c, s.bvh, &info), "Info");
  return info;
}

const trace_bvh* BVH::Handle() const { return s_ ? s_->bvh : nullptr; }
const trace_mesh* BVH::MeshHandle() const { return s_ ? s_->mesh->mesh : nullptr; }

std::vector<Hit> BVH::ClosestHit(const std::vector<Ray>& rays, double tMin, double tMax) const {
  const State& s = state_of(s_, "ClosestHit");
  const size_t n = rays.size();
  std::vector<Hit> out(n);
  if (n == 0) return out;
  std::vector<float> r(8 * n);
  for (size_t i = 0; i < n; ++i) {
    const Ray& q = rays[i];
    float* d = &r[8 * i];
    d[0] = float(q.Origin.X), d[1] = float(q.Origin.Y), d[2] = float(q.Origin.Z), d[3] = float(tMin);
    d[4] = float(q.Direction.X), d[5] = float(q.Direction.Y), d[6] = float(q.Direction.Z), d[7] = float(tMax);
  }
  std::vector<int32_t> hit(n);
  std::vector<float> t(n);
  {
    Lock l(g_mu);
    trace_ctx* c = ctx_locked();
    check(c, trace_closest_hit(c, s.bvh, r.data(), int64_t(n), hit.data(), t.data()), "ClosestHit");
  }
  for (size_t i = 0; i < n; ++i) out[i] = Hit{hit[i] >= 0, hit[i], double(t[i])};
  return out;
}

std::vector<int32_t> BVH::TraversalSteps(const std::vector<Ray>& rays, double tMin, double tMax) const {
  const State& s = state_of(s_, "TraversalSteps");
  const size_t n = rays.size();
  std::vector<int32_t> out(n);
  if (n == 0) return out;
  std::vector<double> r(8 * n);
  for (size_t i = 0; i < n; ++i) {
    const Ray& q = rays[i];
    double* d = &r[8 * i];
    d[0] = q.Origin.X, d[1] = q.Origin.Y, d[2] = q.Origin.Z, d[3] = tMin;
    d[4] = q.Direction.X, d[5] = q.Direction.Y, d[6] = q.Direction.Z, d[7] = tMax;
  }
  Lock l(g_mu);
  trace_ctx* c = ctx_locked();
  check(c, trace_traversal_steps(c, s.bvh, r.data(), int64_t(n), out.data()), "TraversalSteps");
  return out;
}

BVH DefaultLBVH(const Tracables& prims) { return lbvh_of(upload(prims), nullptr); }

BVH LBVH(const Tracables& prims, const AABB& enclosing, int /*threads*/) {
  const double box[6] = {enclosing.Min.X, enclosing.Min.Y, enclosing.Min.Z,
                         enclosing.Max.X, enclosing.Max.Y, enclosing.Max.Z};
  return lbvh_of(upload(prims), box);
}

// ---- PHR -------------------------------------------------------------------------------

PhrBuilder NewPHRBuilder(const Tracables& prims, double alpha, double delta, int branchingFactor,
                         int threadCount) {
  PhrBuilder b;
  b.Alpha = alpha;
  b.Delta = delta;
  b.BranchingFactor = branchingFactor;
  b.Primitives = prims;
  b.ThreadCount = threadCount;
  return b;
}

PhrBuilder NewDefaultBuilder(const Tracables& prims) { return NewPHRBuilder(prims, 0.5, 6, 2, 0); }

double DefaultThreshold(double surface, double alpha, double delta, int depth) {
  return surface / std::pow(2.0, alpha * double(depth) + delta);
}

BVH PhrBuilder::Build() const { return BuildFromAuxilary(DefaultLBVH(Primitives)); }

BVH PhrBuilder::BuildFromAuxilary(const BVH& aux) const { return BuildWithCost(aux).first; }

std::pair<BVH, int64_t> PhrBuilder::BuildWithCost(const BVH& aux) const {
  const std::shared_ptr<BVH::State>& a = BvhAccess::of(aux);
  state_of(a, "BuildFromAuxilary");
  auto st = std::make_shared<BVH::State>();
  st->mesh = a->mesh;  // the PHR tree indexes the auxiliary tree's primitives
  int64_t cost = 0;
  Lock l(g_mu);
  trace_ctx* c = ctx_locked();
  check(c, trace_phr_build_with_cost(c, a->bvh, Alpha, Delta, BranchingFactor, MaxCutSize, &st->bvh, &cost),
        "PHR build");
  return {BVH(st), cost};
}

// ---- PHR parameter search ----------------------------------------------------------------

GridOptimizer NewGridOptimizer(std::vector<double> alphas, std::vector<double> deltas, int frameWidth,
                               int frameHeight) {
  GridOptimizer g;
  g.Alphas = std::move(alphas);
  g.Deltas = std::move(deltas);
  g.FrameWidth = frameWidth;
  g.FrameHeight = frameHeight;
  return g;
}

GridOptimizer NewDefaultGridOptimizer(int frameWidth, int frameHeight) {
  return NewGridOptimizer({0.4, 0.45, 0.5, 0.55}, {6, 7, 8, 9}, frameWidth, frameHeight);
}

std::pair<double, double> GridOptimizer::OptimizedPHRparams(const BVH& aux, int branching, int /*threads*/) const {
  const BVH::State& a = state_of(BvhAccess::of(aux), "OptimizedPHRparams");
  double alpha = 0, delta = 0;
  Lock l(g_mu);
  trace_ctx* c = ctx_locked();
  check(c,
        trace_phr_optimize_grid(c, a.bvh, Alphas.data(), int32_t(Alphas.size()), Deltas.data(),
                                int32_t(Deltas.size()), branching, 2500, FrameWidth, FrameHeight, nullptr, nullptr,
                                &alpha, &delta),
        "grid search");
  return {alpha, delta};
}

BayesianOptimizer NewBayesianOptimizer(std::array<double, 2> alphaRange, std::array<double, 2> deltaRange,
                                       int frameWidth, int frameHeight) {
  BayesianOptimizer b;
  b.AlphaRange = alphaRange;
  b.DeltaRange = deltaRange;
  b.FrameWidth = frameWidth;
  b.FrameHeight = frameHeight;
  return b;
}

BayesianOptimizer NewDefaultBayesianOptimizer(int frameWidth, int frameHeight) {
  return NewBayesianOptimizer({0.4, 0.55}, {6, 9}, frameWidth, frameHeight);
}

std::pair<double, double> BayesianOptimizer::OptimizedPHRparams(const BVH& aux, int branching,
                                                                int /*threads*/) const {
  const BVH::State& a = state_of(BvhAccess::of(aux), "OptimizedPHRparams");
  double alpha = 0, delta = 0;
  Lock l(g_mu);
  trace_ctx* c = ctx_locked();
  check(c,
        trace_phr_optimize_bayes(c, a.bvh, AlphaRange.data(), DeltaRange.data(), branching, 2500, FrameWidth,
                                 FrameHeight, Rounds, RandomRounds, Seed, nullptr, &alpha, &delta),
        "Bayesian search");
  return {alpha, delta};
}

// ---- camera ---------------------------------------------------------------------------------

Camera::Camera(double aspectRatio, double fov, const CameraTransformation& t) {
  viewportHeight_ = 2.0 * std::tan(DegreesToRadians(fov) / 2);
  viewportWidth_ = aspectRatio * viewportHeight_;
  SetTransformation(t);
}

void Camera::corner() {
  lowerLeft_ = Sub(Sub(Sub(origin_, Mul(horizontal_, 0.5)), Mul(vertical_, 0.5)), w_);
}

void Camera::SetTransformation(const CameraTransformation& t) {
  w_ = Unit(Sub(t.LookFrom, t.LookAt));
  u_ = Unit(Cross(t.Up, w_));
  v_ = Cross(w_, u_);
  origin_ = t.LookFrom;
  up_ = t.Up;
  horizontal_ = Mul(u_, viewportWidth_);
  vertical_ = Mul(v_, viewportHeight_);
  corner();
}

void Camera::Translate(Vector3 v) {
  origin_ = Add(origin_, v);
  corner();
}

void Camera::SetFront(Vector3 v) {
  w_ = Unit(v);
  u_ = Unit(Cross(up_, w_));
  v_ = Cross(w_, u_);
  horizontal_ = Mul(u_, viewportWidth_);
  vertical_ = Mul(v_, viewportHeight_);
  corner();
}

trace_camera Camera::Raw() const {
  trace_camera c{};
  const Vector3* src[4] = {&origin_, &lowerLeft_, &horizontal_, &vertical_};
  double* dst[4] = {c.origin, c.lower_left, c.horizontal, c.vertical};
  for (int k = 0; k < 4; ++k) dst[k][0] = src[k]->X, dst[k][1] = src[k]->Y, dst[k][2] = src[k]->Z;
  return c;
}

std::shared_ptr<Camera> NewCamera(double aspectRatio, double fov, const CameraTransformation& t) {
  return std::make_shared<Camera>(aspectRatio, fov, t);
}

std::shared_ptr<Camera> NewDefaultCamera(double aspectRatio, double fov) {
  return NewCamera(aspectRatio, fov, CameraTransformation{Vector3{0, 0, 0}, Vector3{1, 0, 0}, Vector3{0, 1, 0}});
}

// ---- buffers ---------------------------------------------------------------------------------

namespace {
void check_size(int width, int height) {
  if (width < 0 || height < 0) fail(TRACE_E_INVALID, "buffer size must not be negative");
}
}  // namespace

PixelBuffer::PixelBuffer(int width, int height) : width_(width), height_(height) {
  check_size(width, height);
  buff_.resize(size_t(width) * size_t(height));
}

void PixelBuffer::AddSamples(const float* sum, const float* /*last*/, int spp) {
  if (spp <= 0) return;
  for (size_t i = 0; i < buff_.size(); ++i) {
    Pixel& px = buff_[i];
    const double n = double(px.Samples), tot = n + double(spp);
    px.Mean = Color{(n * px.Mean.X + double(sum[3 * i])) / tot, (n * px.Mean.Y + double(sum[3 * i + 1])) / tot,
                    (n * px.Mean.Z + double(sum[3 * i + 2])) / tot};
    px.Samples += spp;
  }
}

FrameBuffer::FrameBuffer(int width, int height) : width_(width), height_(height) {
  check_size(width, height);
  buff_.resize(size_t(width) * size_t(height));
}

void FrameBuffer::AddSamples(const float* /*sum*/, const float* last, int spp) {
  if (spp <= 0) return;
  for (size_t i = 0; i < buff_.size(); ++i)
    buff_[i] = Color{double(last[3 * i]), double(last[3 * i + 1]), double(last[3 * i + 2])};
}

std::unique_ptr<PixelBuffer> NewPxlBuffer(int width, int height) {
  return std::make_unique<PixelBuffer>(width, height);
}
std::unique_ptr<PixelBuffer> NewPxlBufferAR(int height, double aspect) {
  return NewPxlBuffer(int(double(height) * aspect), height);
}
std::unique_ptr<FrameBuffer> NewFrameBuffer(int width, int height) {
  return std::make_unique<FrameBuffer>(width, height);
}
std::unique_ptr<FrameBuffer> NewFrameBufferAR(int height, double aspect) {
  return NewFrameBuffer(int(double(height) * aspect), height);
}

// ---- renderers -----------------------------------------------------------------------------

void ImageRenderer::RenderToBuffer(Buffer& buff) {
  const BVH::State& s = state_of(BvhAccess::of(Bvh), "RenderToBuffer");
  if (!Camera) fail(TRACE_E_INVALID, "RenderToBuffer: no camera");
  if (Verbose) std::fprintf(stderr, "Started rendering\n");
  const int w = buff.Width(), h = buff.Height();
  if (Spp > 0 && w > 0 && h > 0) {
    const size_t n = size_t(w) * size_t(h) * 3;
    std::vector<float> sum(n), sumsq(n), last(n);
    trace_render_params prm{};
    prm.width = w;
    prm.height = h;
    prm.spp = Spp;
    prm.max_depth = MaxDepth;
    prm.miss = int32_t(Miss);
    prm.seed = Seed;
    prm.sample_offset = drawn_;
    const trace_camera cam = Camera->Raw();
    {
      Lock l(g_mu);
      trace_ctx* c = ctx_locked();
      check(c, trace_render(c, s.mesh->mesh, s.bvh, &cam, &prm, sum.data(), sumsq.data(), last.data()),
            "RenderToBuffer");
    }
    buff.AddSamples(sum.data(), last.data(), Spp);
    drawn_ += Spp;
  }
  if (Verbose) std::fprintf(stderr, "Finished Rendering\n");
}

namespace {
std::unique_ptr<ImageRenderer> renderer(const BVH& bvh, std::shared_ptr<Camera> camera, int spp, MissShader miss) {
  auto r = std::make_unique<ImageRenderer>();
  r->MaxDepth = 5;
  r->Spp = spp;
  r->Bvh = bvh;
  r->Camera = std::move(camera);
  r->Miss = miss;
  return r;
}
}  // namespace

std::unique_ptr<ImageRenderer> NewDefaultRenderer(const BVH& bvh, std::shared_ptr<Camera> camera) {
  return renderer(bvh, std::move(camera), 300, DefaultMissShader);
}
std::unique_ptr<ImageRenderer> NewRealtimeRenderer(const BVH& bvh, std::shared_ptr<Camera> camera) {
  return renderer(bvh, std::move(camera), 1, SkyMissShader);
}
std::unique_ptr<ImageRenderer> NewBenchmarkRenderer(const BVH& bvh, std::shared_ptr<Camera> camera) {
  return renderer(bvh, std::move(camera), 1, WhiteMissShader);
}

void HeatMapRenderer::RenderToBuffer(Buffer& buff) {
  const BVH::State& s = state_of(BvhAccess::of(Bvh), "RenderToBuffer");
  if (!Camera) fail(TRACE_E_INVALID, "RenderToBuffer: no camera");
  const int w = buff.Width(), h = buff.Height();
  if (w <= 0 || h <= 0) return;
  std::vector<float> rgb(size_t(w) * size_t(h) * 3);
  const trace_camera cam = Camera->Raw();
  {
    Lock l(g_mu);
    trace_ctx* c = ctx_locked();
    check(c, trace_render_heatmap(c, s.bvh, &cam, w, h, Threshold, rgb.data(), nullptr), "RenderToBuffer");
  }
  buff.AddSamples(rgb.data(), rgb.data(), 1);
}

std::unique_ptr<HeatMapRenderer> NewHeatMapRenderer(const BVH& bvh, std::shared_ptr<Camera> camera,
                                                    int threshold) {
  auto r = std::make_unique<HeatMapRenderer>();
  r->Bvh = bvh;
  r->Camera = std::move(camera);
  r->Threshold = threshold;
  return r;
}

Color DefaultTraversalCountShader(int threshold, int count) {
  if (count > threshold) return Color{double(count) / double(threshold) * 2, 0, 0};
  const double f = double(count) / (double(threshold) * 1.25);
  return Color{0, f, 1 - f};
}

// ---- scene graph -------------------------------------------------------------------------------

Tracables Mesh::raw() const {
  Tracables t;
  t.Prims = geometry;
  t.MaterialIndex.assign(geometry.size(), 0);
  t.Materials.push_back(material);
  return t;
}

std::shared_ptr<Mesh> NewMesh(Geometry geometry, Material mat) {
  auto m = std::make_shared<Mesh>();
  m->geometry = std::move(geometry);
  m->material = mat;
  return m;
}

SceneNode::SceneNode(std::shared_ptr<Mesh> mesh) : transformation_(IdentityMatrix()), mesh_(std::move(mesh)) {}
void SceneNode::Add(std::shared_ptr<SceneNode> node) {
  if (!node) fail(TRACE_E_INVALID, "SceneNode.Add: null node");
  children_.push_back(std::move(node));
}
void SceneNode::ScaleUniform(double factor) { transformation_ = MultiplyMatrix(transformation_, pt::ScaleUniform(factor)); }
void SceneNode::Scale(double x, double y, double z) { transformation_ = MultiplyMatrix(transformation_, pt::Scale(x, y, z)); }
void SceneNode::Translate(double x, double y, double z) {
  transformation_ = MultiplyMatrix(transformation_, pt::Translate(x, y, z));
}
void SceneNode::Rotate(Vector3 dir, double angle) { transformation_ = MultiplyMatrix(transformation_, pt::Rotate(dir, angle)); }

std::shared_ptr<SceneNode> NewSceneNode(std::shared_ptr<Mesh> mesh) { return std::make_shared<SceneNode>(std::move(mesh)); }

Scene::Scene() : root_(NewSceneNode(nullptr)) {}
std::shared_ptr<Scene> NewScene() { return std::make_shared<Scene>(); }

namespace {

void collect_raw(const SceneNode& n, Tracables& out) {
  if (n.MeshOf()) out.Append(n.MeshOf()->raw());
  for (const auto& c : n.Children()) collect_raw(*c, out);
}

// The node arrays trace_scene_compile takes: depth first below the root, each node
// after its parent, siblings in order; meshes listed once each, in first-use order.
struct FlatGraph {
  std::vector<int32_t> parent, node_mesh;
  std::vector<double> matrix;
  std::vector<const Mesh*> meshes;
  std::unordered_map<const Mesh*, int32_t> mesh_index;
};

void flatten(const SceneNode& n, int32_t parent, FlatGraph& g) {
  const int32_t self = int32_t(g.parent.size());
  g.parent.push_back(parent);
  int32_t mi = -1;
  if (n.MeshOf() && !n.MeshOf()->geometry.empty()) {
    const Mesh* m = n.MeshOf().get();
    const auto it = g.mesh_index.emplace(m, int32_t(g.meshes.size())).first;
    mi = it->second;
    if (mi == int32_t(g.meshes.size())) g.meshes.push_back(m);
  }
  g.node_mesh.push_back(mi);
  g.matrix.insert(g.matrix.end(), n.Transformation().begin(), n.Transformation().end());
  for (const auto& c : n.Children()) flatten(*c, self, g);
}

}  // namespace

Tracables Scene::UntransformedTracables() const {
  Tracables t;
  collect_raw(*root_, t);
  return t;
}

namespace {

// The scene's tracables, flattened on the device into a shaded mesh.
std::shared_ptr<MeshState> compile_scene(const SceneNode& root) {
  FlatGraph g;
  // the root itself carries the identity and no mesh: its children are the top level
  for (const auto& c : root.Children()) flatten(*c, -1, g);
  if (root.MeshOf() && !root.MeshOf()->geometry.empty()) fail(TRACE_E_INVALID, "Scene: the root holds no mesh");
  std::vector<int64_t> off{0};
  std::vector<double> pos, nrm;
  std::vector<int32_t> mesh_mat;
  std::vector<float> mats4;
  std::vector<Material> mats;
  bool specular = false;
  for (size_t k = 0; k < g.meshes.size(); ++k) {
    const Tracables t = g.meshes[k]->raw();
    std::vector<double> p, q;
    pack(t, p, q);
    pos.insert(pos.end(), p.begin(), p.end());
    nrm.insert(nrm.end(), q.begin(), q.end());
    off.push_back(off.back() + int64_t(t.size()));
    const Material& m = g.meshes[k]->material;
    mats.push_back(m);
    specular = specular || m.Kind == TRACE_MAT_REFLECTIVE || m.Kind == TRACE_MAT_REFRACTIVE;
    // the scene upload takes Diffuse and Light; the other kinds are set below
    const bool plain = m.Kind == TRACE_MAT_DIFFUSE || m.Kind == TRACE_MAT_LIGHT;
    mats4.insert(mats4.end(), {float(plain ? m.Kind : TRACE_MAT_DIFFUSE), float(m.Albedo.X), float(m.Albedo.Y),
                               float(m.Albedo.Z)});
    mesh_mat.push_back(int32_t(k));
  }
  // the flattened count: a mesh is uploaded once but placed by every node that holds it
  int64_t flat = 0;
  for (int32_t mi : g.node_mesh)
    if (mi >= 0) flat += int64_t(g.meshes[size_t(mi)]->geometry.size());
  if (flat == 0) fail(TRACE_E_INVALID, "Scene: no primitives to compile");
  auto ms = std::make_shared<MeshState>();
  Lock l(g_mu);
  trace_ctx* c = ctx_locked();
  trace_scene* sc = nullptr;
  check(c,
        trace_scene_create(c, pos.data(), nrm.data(), off.data(), int32_t(g.meshes.size()), mesh_mat.data(),
                           mats4.data(), int32_t(g.meshes.size()), &sc),
        "scene upload");
  const int rc = trace_scene_compile(c, sc, g.parent.data(), g.node_mesh.data(), g.matrix.data(),
                                     int32_t(g.parent.size()), &ms->mesh);
  trace_scene_destroy(sc);
  check(c, rc, "scene compile");
  ms->n = flat;
  if (specular) {
    // the compiled FP32 normals and material indices, re-set with the full table
    // (FP32 values are exact in FP64, so the renderer sees the same normals)
    std::vector<float> n32(size_t(9) * size_t(ms->n));
    std::vector<int32_t> idx(size_t(ms->n));
    check(c, trace_mesh_export_shading(c, ms->mesh, n32.data(), idx.data()), "scene shading export");
    const std::vector<double> n64(n32.begin(), n32.end());
    const std::vector<float> mats8 = material_table8(mats);
    check(c, trace_mesh_set_shading_ex(c, ms->mesh, n64.data(), idx.data(), mats8.data(), int32_t(mats.size())),
          "scene shading");
  }
  return ms;
}

}  // namespace

BVH Scene::CompileLBVH() { return lbvh_of(compile_scene(*root_), nullptr); }

BVH Scene::CompilePHR(double alpha, double delta, int branchingFactor) {
  PhrBuilder b = NewPHRBuilder(Tracables{}, alpha, delta, branchingFactor, 0);
  return b.BuildFromAuxilary(CompileLBVH());
}

BVH Scene::Compile() { return CompilePHR(0.5, 6, 2); }

}  // namespace pt
