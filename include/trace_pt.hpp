// trace_pt.hpp -- the reference's pkg/pt API for the rendering hot path, in C++,
// over the C-ABI of trace_hip.h (libtrace_pt.so, linked against libtrace_hip.so).
//
// The reference is Go (pkg/pt); this is the host side a C++ caller uses in its
// place, with the reference's names, argument meanings and call order: build a
// scene or a list of tracables, DefaultLBVH / NewPHRBuilder(...).Build(), a
// camera and a buffer, NewDefaultRenderer(...).RenderToBuffer(buffer).  Every
// build, query and render runs on the GPU; there is no CPU path (a process
// without a device gets pt::Error with code TRACE_E_NODEVICE from the first
// call that needs one, where the Go shim keeps the reference's CPU code).
//
// Where the reference's behaviour cannot carry over, the difference is stated
// at the declaration: threads arguments are accepted and ignored (the device
// decides its parallelism); renders draw their samples from a counter-based
// stream keyed by ImageRenderer::Seed instead of a time-seeded math/rand, so a
// renderer's calls are reproducible and successive calls draw new samples;
// PhrBuilder has no Threshold / Split fields (the device builds DefaultThreshold
// with SweepSAH only); ClosestHitShader and Sampling are fixed to
// DefaultClosestHitShader and RandomSampling.
//
// Errors: a failed call throws pt::Error, whose Code() is the TRACE_E_* status
// and whose what() is the library's message (the reference panics).
//
// One device context per process (device 0 unless SetDevice is called first),
// created on first use; calls on it are serialised by a process-wide lock.
#pragma once

#include <array>
#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "trace_hip.h"

namespace pt {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int Code() const { return code_; }

 private:
  int code_;
};

// Selects the device of the process-wide context; only before its first use
// (TRACE_E_INVALID afterwards).
void SetDevice(int device);
// The process-wide context (created on first use).  For callers that mix this
// API with direct trace_hip.h calls; such calls must hold no other context's
// handles.
trace_ctx* DeviceContext();

// ---- math (math.go:47-100, 144-162, 299-322; transformation.go; util.go:42-44) ----
// FP64, every operation rounded on its own, in the reference's operation order.
struct Vector3 {
  double X = 0, Y = 0, Z = 0;
};
inline Vector3 NewVector3(double x, double y, double z) { return Vector3{x, y, z}; }
using Color = Vector3;
inline Color NewColor(double r, double g, double b) { return Color{r, g, b}; }

Vector3 Add(Vector3 a, Vector3 b);
Vector3 Sub(Vector3 a, Vector3 b);
Vector3 Mul(Vector3 a, double s);
double Dot(Vector3 a, Vector3 b);
Vector3 Cross(Vector3 a, Vector3 b);
double Length(Vector3 a);
Vector3 Unit(Vector3 a);

// Row-major 4x4 (math.go:144): element (r, c) at [4 r + c]; points are columns.
using Matrix4 = std::array<double, 16>;
Matrix4 IdentityMatrix();
// a.MultiplyMatrix(b) (math.go:299-322): the product a b.
Matrix4 MultiplyMatrix(const Matrix4& a, const Matrix4& b);
Matrix4 ScaleUniform(double factor);
Matrix4 Scale(double x, double y, double z);
Matrix4 Translate(double x, double y, double z);
// Rotation by `angle` radians about `dir` through the unit quaternion
// (cos(angle/2), sin(angle/2) dir/|dir|) (transformation.go:27-33, math.go:341-353).
// cos and sin are the C library's, which may differ from Go's math.Cos / math.Sin
// in the last bit.
Matrix4 Rotate(Vector3 dir, double angle);
double DegreesToRadians(double degree);

// ---- materials (material.go:13-106) ----------------------------------------------
// One material value; the four reference types below convert to it.
struct Material {
  int32_t Kind = TRACE_MAT_DIFFUSE;  // TRACE_MAT_*
  Color Albedo;                      // Diffuse/Reflective/Refractive Albedo, Light Color
  double Param = 0;                  // Reflective.Diffusion or Refractive.Ratio
};
struct Diffuse {
  Color Albedo;
  operator Material() const { return Material{TRACE_MAT_DIFFUSE, Albedo, 0}; }
};
struct Light {
  pt::Color Color;
  operator Material() const { return Material{TRACE_MAT_LIGHT, Color, 0}; }
};
struct Reflective {
  Color Albedo;
  double Diffusion = 0;  // in [0, 1]
  operator Material() const { return Material{TRACE_MAT_REFLECTIVE, Albedo, Diffusion}; }
};
struct Refractive {
  Color Albedo;
  double Ratio = 1;
  operator Material() const { return Material{TRACE_MAT_REFRACTIVE, Albedo, Ratio}; }
};

// ---- primitives (primitive.go:25-138) and OBJ files (obj.go) ------------------------
struct Vertex {
  Vector3 Position;
  Vector3 Normal;
};
// A Triangle (three vertices with their normals) or a Sphere (centre, radius).
struct Primitive {
  bool IsSphere = false;
  std::array<Vertex, 3> Vertices{};
  Vector3 Center;
  double Radius = 0;
};
Primitive NewTriangle(const std::array<Vertex, 3>& vertices);
// Normals as the reference computes them (calcNormal, primitive.go:116-138).
Primitive NewTriangleWithoutNormals(Vector3 v0, Vector3 v1, Vector3 v2);
Primitive NewSphere(Vector3 center, double radius);
using Geometry = std::vector<Primitive>;
// parseOBJ (obj.go:10-140): v, vn and f records (f as v, v/vt or v/vt/vn, negative
// indices from the end, polygons as fans from their first vertex; vertex normals
// only when every corner names one); other records are skipped.  Throws pt::Error
// (TRACE_E_INVALID) where the reference panics.
Geometry ParseFromPath(const std::string& path);

// []tracable: primitives with their materials, in scene order.
struct Tracables {
  std::vector<Primitive> Prims;
  std::vector<int32_t> MaterialIndex;  // per primitive, into Materials
  std::vector<Material> Materials;
  size_t size() const { return Prims.size(); }
  void Append(const Tracables& other);
};

struct AABB {
  Vector3 Min, Max;
};
// enclosing (bounding.go:19-25) of the tracables' boxes.
AABB Enclosing(const Tracables& prims);

// ---- rays and hits -------------------------------------------------------------------
struct Ray {
  Vector3 Origin;
  Vector3 Direction;  // need not be unit length
};
struct Hit {
  bool Ok = false;    // false: the ray hit nothing in [tMin, tMax]
  int64_t Index = -1; // the primitive's index in the tracables the tree was built from
  double T = 0;       // hit distance (FP32 precision)
};

// ---- BVH (bvh.go, lbvh.go, phr.go) -------------------------------------------------
class BVH {
 public:
  BVH() = default;
  explicit operator bool() const { return static_cast<bool>(s_); }
  double Cost() const;  // bvh.go:46-48
  int64_t Size() const; // bvh.go:50-52
  // Bulk BVH.intersected (bvh.go:74-77): the closest primitive of each ray with t in
  // [tMin, tMax] (rays and bounds rounded to FP32 for the walk, the winner's t
  // re-solved in FP64).
  std::vector<Hit> ClosestHit(const std::vector<Ray>& rays, double tMin = 0.001,
                              double tMax = std::numeric_limits<double>::infinity()) const;
  // traversalSteps (bvh.go:21-44) of each ray, counted as the reference counts them.
  std::vector<int32_t> TraversalSteps(const std::vector<Ray>& rays, double tMin, double tMax) const;
  trace_bvh_info Info() const;
  const trace_bvh* Handle() const;
  const trace_mesh* MeshHandle() const;

  struct State;
  explicit BVH(std::shared_ptr<State> s) : s_(std::move(s)) {}

 private:
  friend struct BvhAccess;
  std::shared_ptr<State> s_;
};

// lbvh.go:13-25.  `threads` is ignored.
BVH DefaultLBVH(const Tracables& prims);
BVH LBVH(const Tracables& prims, const AABB& enclosing, int threads);

// phr.go:11-84, 154-255.  Threshold and Split are the reference's defaults
// (DefaultThreshold, SweepSAH); MaxCutSize is MAX_CUT_SIZE.
struct PhrBuilder {
  double Alpha = 0.5;
  double Delta = 6;
  int BranchingFactor = 2;  // 2 to 8
  int MaxCutSize = 2500;
  Tracables Primitives;
  int ThreadCount = 0;  // ignored

  // LBVH of Primitives over their enclosing box, then BuildFromAuxilary.
  BVH Build() const;
  BVH BuildFromAuxilary(const BVH& aux) const;
  // The tree and the build cost (the initial cut's size plus every cut handed to
  // SweepSAH).
  std::pair<BVH, int64_t> BuildWithCost(const BVH& aux) const;
};
PhrBuilder NewDefaultBuilder(const Tracables& prims);  // alpha 0.5, delta 6, branching 2
PhrBuilder NewPHRBuilder(const Tracables& prims, double alpha, double delta, int branchingFactor,
                         int threadCount);
// DefaultThreshold (phr.go:356-358): surface / 2^(alpha depth + delta).
double DefaultThreshold(double surface, double alpha, double delta, int depth);

// ---- PHR parameter search (optimizer.go) ---------------------------------------------
struct GridOptimizer {
  std::vector<double> Alphas, Deltas;
  int FrameWidth = 0, FrameHeight = 0;
  // optimizer.go:44-94; `threads` is ignored (candidates run on up to four device
  // streams).
  std::pair<double, double> OptimizedPHRparams(const BVH& aux, int branching, int threads) const;
};
GridOptimizer NewDefaultGridOptimizer(int frameWidth, int frameHeight);
GridOptimizer NewGridOptimizer(std::vector<double> alphas, std::vector<double> deltas, int frameWidth,
                               int frameHeight);

// optimizer.go:96-155 with the Gaussian-process search of trace_bayes_minimize in
// place of the reference's third-party optimiser: Rounds evaluations, the first
// RandomRounds uniform, from a stream seeded with Seed.
struct BayesianOptimizer {
  std::array<double, 2> AlphaRange{}, DeltaRange{};
  int FrameWidth = 0, FrameHeight = 0;
  int Rounds = 20, RandomRounds = 5;
  uint64_t Seed = 1;
  std::pair<double, double> OptimizedPHRparams(const BVH& aux, int branching, int threads) const;
};
BayesianOptimizer NewDefaultBayesianOptimizer(int frameWidth, int frameHeight);
BayesianOptimizer NewBayesianOptimizer(std::array<double, 2> alphaRange, std::array<double, 2> deltaRange,
                                       int frameWidth, int frameHeight);

// ---- camera (camera.go:7-97) -----------------------------------------------------------
struct CameraTransformation {
  Vector3 LookFrom;
  Vector3 LookAt;
  Vector3 Up;
};
class Camera {
 public:
  Camera(double aspectRatio, double fov, const CameraTransformation& t);
  void SetTransformation(const CameraTransformation& t);
  void Translate(Vector3 v);
  void SetFront(Vector3 v);
  Vector3 Origin() const { return origin_; }
  Vector3 Up() const { return up_; }
  Vector3 W() const { return w_; }
  // The camera as the device takes it (origin, lowerLeftCorner, horizontal, vertical).
  trace_camera Raw() const;

 private:
  void corner();
  Vector3 origin_, up_, w_, u_, v_;
  double viewportWidth_ = 0, viewportHeight_ = 0;
  Vector3 lowerLeft_, horizontal_, vertical_;
};
// NewDefaultCamera looks from the origin along +x with +y up.  `fov` is vertical, in
// degrees (tan is the C library's).
std::shared_ptr<Camera> NewDefaultCamera(double aspectRatio, double fov);
std::shared_ptr<Camera> NewCamera(double aspectRatio, double fov, const CameraTransformation& t);

// ---- buffers (buffer.go) ------------------------------------------------------------------
// Pixel (x, y) is at index y * Width + x, row 0 at the bottom, as the reference indexes.
class Buffer {
 public:
  virtual ~Buffer() = default;
  virtual int Width() const = 0;
  virtual int Height() const = 0;
  // Folds one device pass into the buffer with the reference's addSample semantics:
  // per pixel the sum of `spp` samples (3 floats) and the last of them.
  virtual void AddSamples(const float* sum, const float* last, int spp) = 0;
};

struct Pixel {
  int64_t Samples = 0;
  Color Mean;
};
// buffer.go:14-69: a running mean per pixel.  A pixel with n samples and mean c that
// receives spp samples of sum S holds n + spp samples of mean (n c + S) / (n + spp),
// what spp calls of c += (s - c) / samples give in exact arithmetic.
class PixelBuffer : public Buffer {
 public:
  PixelBuffer(int width, int height);
  int Width() const override { return width_; }
  int Height() const override { return height_; }
  void AddSamples(const float* sum, const float* last, int spp) override;
  const Pixel& At(int x, int y) const { return buff_[size_t(y) * width_ + x]; }

 private:
  int width_, height_;
  std::vector<Pixel> buff_;
};
// buffer.go:71-110: each sample overwrites, so a pixel holds the last pass's.
class FrameBuffer : public Buffer {
 public:
  FrameBuffer(int width, int height);
  int Width() const override { return width_; }
  int Height() const override { return height_; }
  void AddSamples(const float* sum, const float* last, int spp) override;
  Color At(int x, int y) const { return buff_[size_t(y) * width_ + x]; }

 private:
  int width_, height_;
  std::vector<Color> buff_;
};
// Width int(height * aspect) as the reference truncates it.
std::unique_ptr<PixelBuffer> NewPxlBuffer(int width, int height);
std::unique_ptr<PixelBuffer> NewPxlBufferAR(int height, double aspect);
std::unique_ptr<FrameBuffer> NewFrameBuffer(int width, int height);
std::unique_ptr<FrameBuffer> NewFrameBufferAR(int height, double aspect);

// ---- renderers (render.go) ------------------------------------------------------------------
enum class MissShader : int32_t {
  Default = TRACE_MISS_DEFAULT,  // DefaultMissShader: black
  White = TRACE_MISS_WHITE,      // WhiteMissShader
  Sky = TRACE_MISS_SKY,          // SkyMissShader: (0.8, 0.8, 0.8) to (0.25, 0.35, 0.5) by the unit direction's y
  Sun = TRACE_MISS_SUN,          // SunMissShader: (1, 0.96, 0.78)
};
constexpr MissShader DefaultMissShader = MissShader::Default;
constexpr MissShader WhiteMissShader = MissShader::White;
constexpr MissShader SkyMissShader = MissShader::Sky;
constexpr MissShader SunMissShader = MissShader::Sun;

// ImageRenderer (render.go:20-129): Spp passes of one sample per pixel, paths of at
// most MaxDepth bounces, DefaultClosestHitShader, RandomSampling.  The tree must come
// from DefaultLBVH / LBVH / PhrBuilder / Scene here (it carries the shading data).
class ImageRenderer {
 public:
  int NumCPU = 0;  // ignored
  int MaxDepth = 5;
  int Spp = 300;
  pt::BVH Bvh;
  std::shared_ptr<pt::Camera> Camera;
  MissShader Miss = DefaultMissShader;
  bool Verbose = false;
  uint64_t Seed = 1;  // key of the sample stream

  void RenderToBuffer(Buffer& buff);
  std::shared_ptr<pt::Camera> GetCamera() const { return Camera; }
  void SetBvh(const pt::BVH& bvh) { Bvh = bvh; }
  pt::BVH GetBvh() const { return Bvh; }
  // Samples per pixel drawn by this renderer's previous calls; the next call draws
  // the samples after them.
  int64_t SamplesDrawn() const { return drawn_; }

 private:
  int64_t drawn_ = 0;
};
// render.go:32-72: Spp 300 / MaxDepth 5 / black miss; Spp 1 / sky; Spp 1 / white.
std::unique_ptr<ImageRenderer> NewDefaultRenderer(const BVH& bvh, std::shared_ptr<Camera> camera);
std::unique_ptr<ImageRenderer> NewRealtimeRenderer(const BVH& bvh, std::shared_ptr<Camera> camera);
std::unique_ptr<ImageRenderer> NewBenchmarkRenderer(const BVH& bvh, std::shared_ptr<Camera> camera);

// HeatMapRenderer (render.go:137-197, 248-258): one ray per pixel through
// (x / (W-1), y / (H-1)), coloured by its traversal steps against Threshold.
class HeatMapRenderer {
 public:
  int NumCPU = 0;  // ignored
  pt::BVH Bvh;
  std::shared_ptr<pt::Camera> Camera;
  int Threshold = 50;

  void RenderToBuffer(Buffer& buff);
  std::shared_ptr<pt::Camera> GetCamera() const { return Camera; }
  void SetBvh(const pt::BVH& bvh) { Bvh = bvh; }
  pt::BVH GetBvh() const { return Bvh; }
};
std::unique_ptr<HeatMapRenderer> NewHeatMapRenderer(const BVH& bvh, std::shared_ptr<Camera> camera,
                                                    int threshold);
// DefaultTraversalCountShader (render.go:250-258).
Color DefaultTraversalCountShader(int threshold, int count);

// ---- scene graph (scene.go) -------------------------------------------------------------------
struct Mesh {
  Geometry geometry;
  Material material;
  Tracables raw() const;  // scene.go:130-139
};
std::shared_ptr<Mesh> NewMesh(Geometry geometry, Material mat);

class SceneNode {
 public:
  explicit SceneNode(std::shared_ptr<Mesh> mesh);
  void Add(std::shared_ptr<SceneNode> node);
  // Each composes on the right: transformation = transformation x M (scene.go:59-86).
  void ScaleUniform(double factor);
  void Scale(double x, double y, double z);
  void Translate(double x, double y, double z);
  void Rotate(Vector3 dir, double angle);
  void SetTransformation(const Matrix4& t) { transformation_ = t; }
  void ResetTransformation() { transformation_ = IdentityMatrix(); }
  const Matrix4& Transformation() const { return transformation_; }
  const std::vector<std::shared_ptr<SceneNode>>& Children() const { return children_; }
  const std::shared_ptr<Mesh>& MeshOf() const { return mesh_; }

 private:
  Matrix4 transformation_;
  std::vector<std::shared_ptr<SceneNode>> children_;
  std::shared_ptr<Mesh> mesh_;
};
std::shared_ptr<SceneNode> NewSceneNode(std::shared_ptr<Mesh> mesh);

// The tracables are flattened on the device (trace_scene_compile): depth first, a
// node's mesh before its children, each mesh transformed by its node's product of
// transformations (positions as points, normals by the inverse transpose) unless its
// own node's transformation is the identity (scene.go:95-114).
class Scene {
 public:
  Scene();
  void Add(std::shared_ptr<SceneNode> node) { root_->Add(std::move(node)); }
  BVH Compile();  // NewDefaultBuilder(prims).Build()
  BVH CompileLBVH();
  BVH CompilePHR(double alpha, double delta, int branchingFactor);
  Tracables UntransformedTracables() const;  // collectTracablesRaw

 private:
  std::shared_ptr<SceneNode> root_;
};
std::shared_ptr<Scene> NewScene();

}  // namespace pt
