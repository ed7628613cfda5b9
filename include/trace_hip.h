/*
 * trace_hip.h -- C-ABI of the MI355X (gfx950) rendering hot path.
 *
 * The reference (ChSchmidt99/trace, pure Go) has no FFI; its seams are Go
 * functions whose bodies a cgo shim replaces with the calls below.  Each entry
 * point names the reference function it stands in for (paths relative to the
 * reference checkout).
 *
 * Conventions
 *   - Every call returns an int status: TRACE_OK (0) or a negative TRACE_E_*;
 *     on failure trace_last_error(ctx) holds a message.  (The reference panics
 *     on errors; the Go shim turns a non-zero status into a panic.)
 *     Host memory exhaustion inside a call (a C++ allocation failure)
 *     is not turned into a status: it terminates the process.
 *   - Plain pointers and sizes only.  Pointers named h_* are host memory, d_*
 *     are device memory.  Host inputs are copied during the call and never
 *     retained.  Host-pointer calls are synchronous; *_device calls are
 *     asynchronous on the context's stream (see trace_ctx_set_stream).
 *   - A context is not thread-safe; distinct contexts may be used concurrently.
 *   - Triangles are 9 doubles each: v0.xyz, v1.xyz, v2.xyz.
 *   - Rays are 8 floats each: origin.xyz, tmin, direction.xyz, tmax.  The
 *     direction is not normalised (the reference never normalises it).
 */
#ifndef TRACE_HIP_H
#define TRACE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct trace_ctx trace_ctx;
typedef struct trace_mesh trace_mesh;
typedef struct trace_bvh trace_bvh;
typedef struct trace_scene trace_scene;

enum {
  TRACE_OK = 0,
  TRACE_E_INVALID = -1,     /* bad argument (null handle, n < 0, size mismatch) */
  TRACE_E_NODEVICE = -2,    /* no HIP device / device index out of range */
  TRACE_E_HIP = -3,         /* a HIP runtime call failed */
  TRACE_E_OOM = -4,         /* device allocation failed */
  TRACE_E_UNSUPPORTED = -5, /* input outside what this build supports */
  TRACE_E_DEVICE = -6       /* a kernel reported an internal error (bounded wait expired, stack overflow) */
};

/* Morton key modes for the LBVH build.
 * REF63: the reference's 63-bit code (21 bits per axis, lbvh.go:64-73,
 *        morton.go:6-20), sorted in full; topology equals the reference's.
 * M30:   code63 >> 33 (10 bits per axis); the reference tree truncated at
 *        bit 33, half the radix passes. */
enum { TRACE_KEY_REF63 = 0, TRACE_KEY_M30 = 1 };

/* Miss shaders (render.go:228-246). */
enum { TRACE_MISS_DEFAULT = 0, TRACE_MISS_WHITE = 1, TRACE_MISS_SKY = 2, TRACE_MISS_SUN = 3 };

/* Material kinds (material.go:13-42). */
enum { TRACE_MAT_DIFFUSE = 0, TRACE_MAT_LIGHT = 1, TRACE_MAT_REFLECTIVE = 2, TRACE_MAT_REFRACTIVE = 3 };

int trace_version(void);

/* ---- context ------------------------------------------------------------ */
int trace_ctx_create(int device, trace_ctx** out);
int trace_ctx_destroy(trace_ctx* ctx);
/* Message of the last failed call on ctx ("" if none).  Valid until the next call. */
const char* trace_last_error(const trace_ctx* ctx);
/* Use an external hipStream_t for all work of ctx (NULL restores the ctx's own stream). */
int trace_ctx_set_stream(trace_ctx* ctx, void* hip_stream);
int trace_ctx_synchronize(trace_ctx* ctx);

/* ---- geometry ----------------------------------------------------------- */
/* Device-resident triangle soup in caller order.  Replaces the marshalling of
 * Scene.collectTracables (scene.go:100-114) + NewTriangle (primitive.go:102-114). */
/* A primitive slot whose fifth value is NaN is an analytic Sphere (NewSphere,
 * primitive.go:25-86): centre (3 values), radius (> 0), NaN, then 4 ignored values.
 * Its box is centre -/+ radius, its hit the reference's quadratic (bounds exclusive),
 * its shading normal (point - centre) / radius; scene compile moves its centre only
 * (Sphere.transformed, primitive.go:42-50), and so does trace_mesh_animate. */
int trace_mesh_create(trace_ctx* ctx, const double* h_pos, int64_t ntri, trace_mesh** out);
/* Same from device memory (9 doubles per triangle), asynchronous. */
int trace_mesh_create_device(trace_ctx* ctx, const double* d_pos, int64_t ntri, trace_mesh** out);
/* Many meshes in one buffer: segment s holds triangles [offsets[s], offsets[s+1]),
 * offsets[0] = 0, offsets[nseg] = ntri, no empty segment.  trace_lbvh_build on such a
 * mesh builds one LBVH per segment in a single pass (a forest): each tree equals the
 * LBVH (lbvh.go:13-25) of its mesh alone.  No reference counterpart: the reference
 * builds one scene at a time. */
int trace_mesh_create_segmented(trace_ctx* ctx, const double* h_pos, int64_t ntri,
                                const int64_t* h_seg_offsets, int32_t nseg, trace_mesh** out);
/* A handle on tree `seg` of a forest.  It aliases the forest's memory (destroy it with
 * trace_bvh_destroy before the forest); triangle indices it reports are local to the
 * segment.  Closest hit, export, Cost and render take views; a forest itself is queried
 * through them. */
int trace_bvh_segment(trace_ctx* ctx, const trace_bvh* forest, int32_t seg, trace_bvh** out);
/* Replace the positions of an existing mesh (same ntri); used for per-frame rebuilds
 * (cmd/dynamic/main.go:113-117).  Geometric shading normals follow the new positions. */
int trace_mesh_update(trace_ctx* ctx, trace_mesh* mesh, const double* h_pos);
int trace_mesh_destroy(trace_mesh* mesh);

/* ---- LBVH build --------------------------------------------------------- */
/* LBVH(prims, enclosing, threads) (lbvh.go:17-25); DefaultLBVH (lbvh.go:13-15)
 * when h_enclosing is NULL.  h_enclosing = {minx,miny,minz,maxx,maxy,maxz}
 * is used only for the Morton quantisation, as in the reference. */
int trace_lbvh_build(trace_ctx* ctx, const trace_mesh* mesh, const double* h_enclosing,
                     int key_mode, trace_bvh** out);
/* Rebuild into an existing LBVH handle, reusing its device buffers (grow-only);
 * asynchronous on the context stream, no device allocation once the buffers fit.
 * The per-frame rebuild of cmd/dynamic/main.go:113-117. */
int trace_lbvh_rebuild(trace_ctx* ctx, const trace_mesh* mesh, const double* h_enclosing,
                       int key_mode, trace_bvh* bvh);
/* trace_lbvh_rebuild that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Its scratch is a fixed array of
 * 64 MiB that the library holds on each device until the process ends (about 1.97M
 * triangles of one mesh at the default key sort); larger meshes are refused with
 * TRACE_E_UNSUPPORTED before anything is enqueued.  The tree's own arrays must
 * already be sized: bvh was built from a mesh of the same shape with
 * trace_lbvh_build (else TRACE_E_INVALID).  h_enclosing is passed to the device by
 * value (no host-memory copy).  Asynchronous on the context stream.  Every call on a
 * device shares that device's array, which keeps the last call's contents: calls
 * that overlap on one device (several contexts or streams) race. */
int trace_lbvh_rebuild_noalloc(trace_ctx* ctx, const trace_mesh* mesh, const double* h_enclosing,
                               int key_mode, trace_bvh* bvh);
int trace_bvh_destroy(trace_bvh* bvh);

typedef struct {
  int64_t n_prims;   /* triangles */
  int64_t n_leaves;  /* leaves (LBVH: unique Morton keys) */
  int64_t n_inner;   /* inner nodes */
  int32_t depth;     /* inner-node levels on the longest root-to-leaf path */
  int32_t key_mode;  /* TRACE_KEY_* or -1 for a PHR tree */
  int32_t arity;     /* maximum children per inner node */
  int32_t pad;
  double root_box[6];
  int64_t phr_initial_cut; /* PHR: size of the initial cut (0 for an LBVH) */
  int32_t phr_levels;      /* PHR: job levels processed */
  int32_t n_segments;      /* trees in a forest (1 for a single tree or a segment view) */
  int64_t phr_build_cost;  /* PHR: BuildWithCost's cost (phr.go:154-255): the initial cut's
                              size plus the size of every cut handed to SweepSAH */
} trace_bvh_info;
int trace_bvh_get_info(trace_ctx* ctx, const trace_bvh* bvh, trace_bvh_info* out);

/* Sorted Morton keys and the sort permutation (sorted position -> triangle index). */
int trace_bvh_export_sorted(trace_ctx* ctx, const trace_bvh* bvh, uint64_t* h_keys, int32_t* h_perm);

/* Canonical preorder serialisation (children in stored order, children[0] first),
 * 4 int32 per node: kind (0 inner, 1 leaf), first, last (leaf: sorted-position range
 * of its triangles, inclusive; inner: covered leaf range), n_children; and 6 floats per
 * node: the node's AABB.  Returns the node count in *n_out; fails with TRACE_E_INVALID
 * if cap is too small. */
int trace_bvh_export_canonical(trace_ctx* ctx, const trace_bvh* bvh, int64_t cap,
                               int32_t* h_rec, float* h_box, int64_t* n_out);

/* BVH.Cost() (bvh.go:46-48, 226-239) and BVH.Size() (bvh.go:50-52, 159-167). */
int trace_bvh_cost(trace_ctx* ctx, const trace_bvh* bvh, double* cost);
int trace_bvh_size(trace_ctx* ctx, const trace_bvh* bvh, int64_t* size);

/* ---- PHR ------------------------------------------------------------------ */
/* PhrBuilder{Alpha, Delta, BranchingFactor}.BuildFromAuxilary(aux) (phr.go:30-84, 92-459)
 * with MAX_CUT_SIZE = max_cut (phr.go:11, 2500 in the reference); aux is an LBVH built
 * by trace_lbvh_build.  NewDefaultBuilder is alpha 0.5, delta 6, branching 2
 * (phr.go:26-28).  The initial cut is taken in FIFO order from the root.  Branching
 * factors 2 to 8 are built (TRACE_E_UNSUPPORTED otherwise): each job splits its largest
 * cut until it holds `branching` cuts (phr.go:98-127).  The result is a new BVH handle
 * (its triangles in PHR leaf order); Cost/Size/export/closest-hit/render/traversal steps
 * accept it, and the export reports each branch's own child count. */
int trace_phr_build(trace_ctx* ctx, const trace_bvh* aux, double alpha, double delta,
                    int branching, int32_t max_cut, trace_bvh** out);
/* trace_phr_build into the storage of an existing PHR tree `phr` (a handle that
 * trace_phr_build returned; any aux, alpha, delta, branching and max_cut): the same
 * tree a fresh build gives, without allocating when the storage is large enough --
 * the reference builds a new tree per BuildFromAuxilary call (phr.go:47-84), which a
 * caller that rebuilds per frame or per optimiser candidate replaces with this.  On an
 * error the tree's contents are undefined until a rebuild succeeds.
 * A binary rebuild (branching 2) sizes the tree for inputs of the same triangle count,
 * and may then be recorded in a graph (trace_graph_begin): a rebuild issued
 * while capturing takes the same alpha, delta and branching 2 as the eager one before
 * it (TRACE_E_INVALID otherwise) and records, with no host round trip, the eager
 * rebuild's job levels plus 4.  A replay whose input needs more levels or more than an
 * eighth more branches, makes a deeper tree than the eager one's depth class (63 or
 * 255), or has an initial cut of one node sets the device error word (reported by the next trace_ctx_synchronize)
 * and leaves the previous tree in place.  Build statistics (initial cut, build cost)
 * stay the eager rebuild's; depth and levels (trace_bvh_get_info) and the export
 * follow the replays. */
int trace_phr_rebuild(trace_ctx* ctx, const trace_bvh* aux, double alpha, double delta, int branching,
                      int32_t max_cut, trace_bvh* phr);
/* PhrBuilder.BuildWithCost (phr.go:154-255): trace_phr_build plus its build cost -- the
 * initial cut's size plus the size of every cut handed to SweepSAH. */
int trace_phr_build_with_cost(trace_ctx* ctx, const trace_bvh* aux, double alpha, double delta,
                              int branching, int32_t max_cut, trace_bvh** out, int64_t* build_cost);
/* gridOptimizer.OptimizedPHRparams (optimizer.go:44-94; NewDefaultGridOptimizer is alphas
 * {0.4, 0.45, 0.5, 0.55} x deltas {6, 7, 8, 9}): BuildWithCost and Cost() for every
 * (alpha, delta), alphas outer, then the (alpha, delta) of the first minimum of
 * build / max build + omega^2 * Cost / max Cost with omega = (prims / 2e6)^2 *
 * (frame_width * frame_height / 2e5).  h_sah and h_build (optional, n_alpha * n_delta)
 * receive every evaluation.  Synchronous: waits for ctx's stream, then evaluates the
 * candidates on up to four worker contexts of the same device (made on first use, kept
 * and destroyed with ctx), each on its own stream and host thread. */
int trace_phr_optimize_grid(trace_ctx* ctx, const trace_bvh* aux, const double* h_alphas, int32_t n_alpha,
                            const double* h_deltas, int32_t n_delta, int branching, int32_t max_cut,
                            int32_t frame_width, int32_t frame_height, double* h_sah, int64_t* h_build,
                            double* alpha_out, double* delta_out);
/* bayesianOptimizer.OptimizedPHRparams (optimizer.go:96-155; NewDefaultBayesianOptimizer
 * is alpha in [0.4, 0.55], delta in [6, 9]).  The normalisers are Cost() of the tree at
 * (alpha_range[0], delta_range[0]) and the build cost at (alpha_range[1], delta_range[1]);
 * each of `rounds` evaluations scores build / max build + omega^2 * Cost / max Cost with
 * omega as in trace_phr_optimize_grid (a zero alpha or delta scores the normalisers
 * themselves, optimizer.go:142-144).  The points come from trace_bayes_minimize with
 * `random_rounds` uniform samples first and `seed` (the reference's third-party search,
 * changkun.de/x/bo v0.0.1, is replaced by that function).  h_trace (optional,
 * (rounds + 2) x 5 doubles) receives per row alpha, delta, Cost, build cost, objective:
 * the two normaliser builds (objective NaN), then the evaluations in order.  Writes the
 * evaluated (alpha, delta) of the smallest objective (first minimum).  Returns
 * TRACE_E_INVALID when either normaliser is zero (the reference's objective is then NaN). */
int trace_phr_optimize_bayes(trace_ctx* ctx, const trace_bvh* aux, const double* alpha_range,
                             const double* delta_range, int branching, int32_t max_cut, int32_t frame_width,
                             int32_t frame_height, int32_t rounds, int32_t random_rounds, uint64_t seed,
                             double* h_trace, double* alpha_out, double* delta_out);
/* Objective for trace_bayes_minimize: writes f(x) to *y; a non-zero return aborts the
 * search and is passed on. */
typedef int (*trace_objective_fn)(void* user, const double* x, double* y);
/* bo.Optimizer.RunSerial as called at optimizer.go:141-151 (minimise, expected
 * improvement), host only (no device, no context): minimises `objective` over the box
 * [lo, hi] (dims 1..8) with `rounds` evaluations.  The first max(1, random_rounds) sample
 * the box uniformly from a splitmix64 stream seeded with `seed`; each later one fits a
 * Gaussian process (Matern 5/2 kernel in unit-cube coordinates, standardised values,
 * length scale of largest marginal likelihood) and evaluates the point of largest
 * expected improvement.  h_x (rounds x dims) and h_y (rounds) receive the evaluations
 * (optional); x_best the evaluated point of smallest value (first minimum).  The same
 * seed and objective give the same sequence.  Returns TRACE_E_INVALID for bad arguments
 * or a NaN value. */
int trace_bayes_minimize(const double* lo, const double* hi, int32_t dims, int32_t rounds, int32_t random_rounds,
                         uint64_t seed, trace_objective_fn objective, void* user, double* h_x, double* h_y,
                         double* x_best);
/* For a PHR tree: per PHR-ordered triangle, its sorted position in the LBVH it was
 * built from (relates PHR leaves to the LBVH order). */
int trace_phr_export_source(trace_ctx* ctx, const trace_bvh* phr, int32_t* h_src);

/* ---- instancing (no reference counterpart) ------------------------------------ */
/* A two-level tree for scenes that place meshes many times.  The reference flattens every
 * placement into one triangle list (Scene.collectTracables, scene.go:100-114); here
 * instance i places bottom-level tree h_blas[h_inst_blas[i]] (the LBVH or PHR tree of a
 * triangle mesh in its own coordinates, or a segment view of a forest) with the affine
 * object-to-world transform h_xform[12 i .. 12 i + 11] (rows (A | t): world = A object + t,
 * A invertible), and its triangles take the flattened indices h_tri_base[i] + (the mesh's
 * own index).  The top level is an LBVH over the instances' world boxes.  trace_closest_hit
 * and trace_closest_hit_device accept the result and report flattened indices; rays are
 * moved into each instance's space in FP64 and the winner is re-solved there in FP64, so a
 * hit agrees with the flattened scene's up to rounding at triangle edges (ties go to the
 * larger flattened index).  Export, Cost and Size describe the top level; renders,
 * traversal steps and PHR refinement refuse it (TRACE_E_UNSUPPORTED).  The bottom-level
 * trees must outlive the result and must not hold spheres. */
int trace_instances_build(trace_ctx* ctx, const trace_bvh* const* h_blas, int32_t n_blas, const int32_t* h_inst_blas,
                          const double* h_xform, const int64_t* h_tri_base, int32_t n_inst, trace_bvh** out);
/* Device memory a tree holds, in bytes (a segment view holds none; an instanced tree's
 * count leaves out its bottom-level trees). */
int trace_bvh_device_bytes(trace_ctx* ctx, const trace_bvh* bvh, int64_t* bytes);

/* ---- closest hit -------------------------------------------------------- */
/* Bulk BVH.intersected (bvh.go:74-77, 179-199 + primitive.go:172-208): per ray the
 * closest triangle with t in [tmin, tmax]; hit = triangle index or -1, t = hit
 * distance (tmax on a miss).  Equal-t ties resolve to the triangle the reference's
 * depth-first order visits last. */
int trace_closest_hit(trace_ctx* ctx, const trace_bvh* bvh, const float* h_rays, int64_t nrays,
                      int32_t* h_hit, float* h_t);
int trace_closest_hit_device(trace_ctx* ctx, const trace_bvh* bvh, const float* d_rays,
                             int64_t nrays, int32_t* d_hit, float* d_t);
/* Counting mode: per ray, inner nodes visited and triangles tested (for byte accounting). */
int trace_closest_hit_counts(trace_ctx* ctx, const trace_bvh* bvh, const float* h_rays,
                             int64_t nrays, int32_t* h_nodes, int32_t* h_tris);

/* ---- path tracing -------------------------------------------------------- */
/* Per-triangle shading: vertex normals (9 doubles per triangle as
 * NewTriangle takes them, primitive.go:102-114; NULL = geometric normals as
 * NewTriangleWithoutNormals computes them, primitive.go:116-138), a material
 * index per triangle, and a material table of 4 floats each: kind
 * (TRACE_MAT_DIFFUSE / TRACE_MAT_LIGHT), r, g, b (Diffuse.Albedo or
 * Light.Color, material.go:13-42).  The reference assigns one Material per
 * Mesh (scene.go:118-139); a per-triangle index expresses any set of meshes. */
int trace_mesh_set_shading(trace_ctx* ctx, trace_mesh* mesh, const double* h_normals,
                           const int32_t* h_mat_index, const float* h_materials, int32_t nmat);

/* Camera as camera.go:34-66 stores it (origin, lowerLeftCorner, horizontal,
 * vertical); the direction of sample (s, t) is LL + s*H + t*V - origin. */
typedef struct {
  double origin[3];
  double lower_left[3];
  double horizontal[3];
  double vertical[3];
} trace_camera;

typedef struct {
  int32_t width;
  int32_t height;
  int32_t spp;           /* ImageRenderer.Spp */
  int32_t max_depth;     /* ImageRenderer.MaxDepth */
  int32_t miss;          /* TRACE_MISS_* (ImageRenderer.Miss) */
  int32_t pad;
  uint64_t seed;         /* Philox key of the sample stream */
  int64_t sample_offset; /* index of the first sample (progressive rendering) */
} trace_render_params;

/* ImageRenderer.RenderToBuffer (render.go:91-129) with DefaultClosestHitShader
 * and RandomSampling.  Outputs 3 floats per pixel, row-major with row 0 at the
 * bottom (y as the reference indexes its buffers): the sum and the sum of
 * squares of the spp samples, and the last sample (NULL to skip; FrameBuffer
 * keeps only the last pass, buffer.go:90-92).  The bvh must be built from mesh. */
int trace_render(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh,
                 const trace_camera* cam, const trace_render_params* prm, float* h_sum,
                 float* h_sumsq, float* h_last);
int trace_render_device(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh,
                        const trace_camera* cam, const trace_render_params* prm, float* d_sum,
                        float* d_sumsq, float* d_last);

/* trace_mesh_set_shading with all four material kinds of material.go:13-106: 8 floats
 * per material, kind, r, g, b (Albedo / Color), then the parameter -- Reflective.Diffusion
 * (>= 0) or Refractive.Ratio (> 0), else ignored -- and 3 reserved (0).  Reflective and
 * Refractive scatter as the reference does (reflect, refract with Schlick's reflectance,
 * material.go:44-106); the draw that chooses reflection is the vertex's own Philox
 * stream.  Renders accept them; trace_render_backward does not. */
int trace_mesh_set_shading_ex(trace_ctx* ctx, trace_mesh* mesh, const double* h_normals,
                              const int32_t* h_mat_index, const float* h_materials8, int32_t nmat);

/* ---- scene compile (scene.go:7-151, primitive.go:152-170) ----------------------- */
/* The meshes of a scene graph (Mesh{geometry, material}, scene.go:116-139), uploaded
 * once: mesh k is triangles [h_mesh_off[k], h_mesh_off[k+1]) of h_pos (9 doubles per
 * triangle) with vertex normals h_nrm (9 doubles per triangle; NULL: geometric normals
 * of the flattened triangles) and material h_mesh_mat[k] of the table h_materials
 * (nmat x (kind, r, g, b), Diffuse or Light). */
int trace_scene_create(trace_ctx* ctx, const double* h_pos, const double* h_nrm, const int64_t* h_mesh_off,
                       int32_t n_mesh, const int32_t* h_mesh_mat, const float* h_materials, int32_t nmat,
                       trace_scene** out);
/* Scene.Compile's collectTracables (scene.go:21-36, 95-114) on the device.  The
 * SceneNode tree below Scene.root (identity, no mesh): node k has parent h_parent[k]
 * (-1: a child of the root; otherwise an earlier node), mesh h_node_mesh[k] (-1: none)
 * and transformation h_matrix[16k .. 16k+15] (Matrix4, row-major, math.go:144-162);
 * siblings keep index order.  Flattened depth first (a node's mesh, then its children)
 * with t = node.transformation x parent's t, positions transformed as points, normals
 * by (t^T)^-1 as points, not renormalised (primitive.go:152-170), and meshes whose own
 * node matrix is the identity copied untransformed, as the reference does.  Writes a
 * mesh with shading data ready for trace_lbvh_build: *io_mesh NULL creates it; an
 * existing mesh of the same triangle count is refilled in place. */
int trace_scene_compile(trace_ctx* ctx, const trace_scene* scene, const int32_t* h_parent,
                        const int32_t* h_node_mesh, const double* h_matrix, int32_t n_nodes,
                        trace_mesh** io_mesh);
int trace_scene_destroy(trace_scene* scene);
/* A mesh's shading data as the renderer holds it: FP32 vertex normals (9 per triangle)
 * and material indices; either output may be NULL. */
int trace_mesh_export_shading(trace_ctx* ctx, const trace_mesh* mesh, float* h_nrm, int32_t* h_mat_idx);

/* ---- traversal cost view (bvh.go:21-44, render.go:137-197, 250-258) ---------- */
/* BVH.traversalSteps (bvh.go:21-44) per ray: the reference's own stack walk (children
 * pushed in stored order, FP64 slab and Moeller-Trumbore under the current best t), so
 * the count -- inner nodes whose box the ray enters -- equals the reference's.  Rays are
 * 8 doubles each (origin, tMin, direction, tMax).  Host buffers, synchronous. */
int trace_traversal_steps(trace_ctx* ctx, const trace_bvh* bvh, const double* h_rays, int64_t n,
                          int32_t* h_counts);
int trace_traversal_steps_device(trace_ctx* ctx, const trace_bvh* bvh, const double* d_rays, int64_t n,
                                 int32_t* d_counts);
/* HeatMapRenderer.RenderToBuffer (render.go:167-197): one camera ray per pixel through
 * (x / (W-1), y / (H-1)), tMin 0.001, tMax MaxFloat64, coloured by
 * DefaultTraversalCountShader (render.go:250-258) with the given threshold.  h_rgb gets
 * W*H*3 floats (pixel y*W + x, the PixelBuffer mean of its one sample); h_counts
 * (optional, W*H) the step counts.  The _device variant is asynchronous. */
int trace_render_heatmap(trace_ctx* ctx, const trace_bvh* bvh, const trace_camera* cam, int32_t width,
                         int32_t height, int32_t threshold, float* h_rgb, int32_t* h_counts);
int trace_render_heatmap_device(trace_ctx* ctx, const trace_bvh* bvh, const trace_camera* cam, int32_t width,
                                int32_t height, int32_t threshold, float* d_rgb, int32_t* d_counts);

/* Path segments (rays traced, every bounce) of the context's last trace_render*,
 * trace_render_batched or trace_render_backward call (the backward counts both of its
 * sweeps).  Waits for the context stream.  No reference counterpart: the reference's
 * renderer keeps no ray count; this gives the rays-per-second figure of a render. */
int trace_render_segments(trace_ctx* ctx, int64_t* out);
/* Counting mode of trace_render (the image is discarded): h_counts[0] = path segments
 * traced, [1] = nodes of the 4-wide traversal tree visited, [2] = primitives tested, over
 * all segments.  Synchronous.  No reference counterpart: the byte count behind the
 * render's bandwidth figure, as trace_closest_hit_counts is for closest hit. */
int trace_render_counts(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh, const trace_camera* cam,
                        const trace_render_params* prm, int64_t* h_counts);

/* ---- differentiation (no reference counterpart) ----------------------------- */
/* Bind per-triangle parameters: d_albedo replaces the colour of Diffuse triangles
 * (Diffuse.Albedo, material.go:25-38) and d_emission that of Light triangles
 * (Light.Color, material.go:13-23); 3 floats per triangle, caller order, device memory.
 * The mesh reads them at every render and backward pass and does not copy them: keep
 * them alive while work that reads them is queued.  NULL unbinds (material colours). */
int trace_mesh_bind_params(trace_ctx* ctx, trace_mesh* mesh, const float* d_albedo,
                           const float* d_emission);
/* Gradient of a scalar loss on the mean image (sum / spp) of trace_render_device with
 * the same arguments, with respect to the per-triangle albedo and emission.
 * d_grad_image = dLoss/d(mean image), W*H*3 floats; d_grad_albedo and d_grad_emission
 * receive 3 floats per triangle (caller order; overwritten).  Path replay: each path is
 * traced once to find where it ends and once more to hand every vertex its share, with
 * the same random numbers as the render, and no path is stored.  Asynchronous. */
int trace_render_backward(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh,
                          const trace_camera* cam, const trace_render_params* prm,
                          const float* d_grad_image, float* d_grad_albedo, float* d_grad_emission);
/* trace_render_device that also keeps, per path, what trace_render_backward's first
 * sweep would compute (where the path ended, and the product of its non-zero albedos
 * and the radiance there): d_keep = W*H*spp*4 floats of device memory (16-byte aligned),
 * written sample-major.  The image is the same, bit for bit, as trace_render_device's.
 * Meshes with Reflective or Refractive materials are rejected, as by the backward.
 * Asynchronous.  (No reference counterpart: for inverse-rendering loops, which render
 * the primal image anyway.) */
int trace_render_keep(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh, const trace_camera* cam,
                      const trace_render_params* prm, float* d_sum, float* d_sumsq, float* d_last,
                      float* d_keep);
/* trace_render_backward given the d_keep of a trace_render_keep call with the same
 * arguments and the same bound parameters: only the replay sweep is traced.  The
 * gradient equals trace_render_backward's (up to the order of its atomic sums).
 * Asynchronous. */
int trace_render_backward_kept(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh,
                               const trace_camera* cam, const trace_render_params* prm, const float* d_keep,
                               const float* d_grad_image, float* d_grad_albedo, float* d_grad_emission);

/* One Adam step (Kingma & Ba 2015) on n device parameters, in place: m, v are the
 * caller's moment buffers (zero before step 1), step counts from 1; the result is
 * clamped to [lo, hi] (albedos stay in [0, 1]).  Asynchronous.  The reference's
 * optimiser (optimizer.go:44-155) searches PHR parameters instead; this drives the
 * inverse-rendering loop over per-triangle albedo. */
int trace_adam_step(trace_ctx* ctx, float* d_param, const float* d_grad, float* d_m, float* d_v,
                    int64_t n, float lr, float beta1, float beta2, float eps, int32_t step, float lo,
                    float hi);

/* ---- batching, graphs, tuning (no reference counterpart) ------------------- */
/* nscene images in one wavefront: scene c traces tree h_seg_ids[c] of a forest (built
 * from mesh, a segmented mesh) with camera h_cams[c] and the Philox stream of
 * prm->seed + c, so it equals a trace_render of that segment's view with seed + c.
 * Device outputs hold nscene * W * H * 3 floats each, scene-major.  Asynchronous. */
int trace_render_batched(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* forest,
                         int32_t nscene, const int32_t* h_seg_ids, const trace_camera* h_cams,
                         const trace_render_params* prm, float* d_sum, float* d_sumsq,
                         float* d_last);

/* Per-frame vertex animation on the device: mesh positions = rest positions moved by
 * amplitude * sin(2*pi*frame/period + k.p) along a unit direction u (k, u drawn from
 * seed; a vertex's displacement depends only on its rest position, so shared vertices
 * stay shared).  frame is read from *d_frame and then incremented, so a captured graph
 * advances one frame per launch.  Geometric shading normals (trace_mesh_set_shading
 * without normals) follow the new positions; caller normals are kept.  The per-frame
 * rebuild of cmd/dynamic/main.go:113-117. */
int trace_mesh_animate(trace_ctx* ctx, trace_mesh* mesh, const trace_mesh* rest,
                       double amplitude, double period, uint64_t seed, uint32_t* d_frame);
/* Copy a mesh's current positions (9 doubles per triangle) to the host. */
int trace_mesh_export(trace_ctx* ctx, const trace_mesh* mesh, double* h_pos);

/* HIP graph capture of the context stream: asynchronous calls issued between begin
 * and end (trace_mesh_animate, trace_lbvh_rebuild, trace_phr_rebuild with branching 2,
 * trace_closest_hit_device, trace_render_device) are recorded once and replayed by
 * trace_graph_launch.  Device
 * buffers must already be sized (run the work once eagerly first); no call that
 * synchronises may be issued while capturing.  While a graph is alive the context's
 * scratch arena cannot grow (a call that needs more fails with TRACE_E_INVALID);
 * destroy graphs before their context. */
typedef struct trace_graph trace_graph;
int trace_graph_begin(trace_ctx* ctx);
int trace_graph_end(trace_ctx* ctx, trace_graph** out);
int trace_graph_launch(trace_ctx* ctx, const trace_graph* graph);
int trace_graph_destroy(trace_graph* graph);

/* Launch parameters chosen by measurement on the device at hand. */
typedef struct {
  int32_t trav_threads;   /* closest-hit workgroup size: 64, 128 or 256 */
  int32_t render_threads; /* path-segment workgroup size: 64, 128 or 256 */
  int64_t render_batch;   /* paths per wavefront batch (default 2^27: <= 9.7 GB of path state) */
  int32_t tuned;          /* 1 once an autotune call has run */
  int32_t trav_persistent; /* closest hit: 1 = persistent lanes refilled from a work counter,
                              0 = one ray per lane */
  int32_t render_persistent; /* path segments: 1 = split (persistent trace pass over the
                                queue, then a shade pass), 0 = fused (one segment per lane) */
  int32_t sort_bits;            /* LBVH key sort: radix digit width, 8, 10 or 11 bits */
  int32_t sort_keys_per_thread; /* LBVH key sort: keys per thread, 4, 8 or 16 (tiles of 256x);
                                   valid pairs: 8-bit digits with 4, 8 or 16; 10-bit with 4 or
                                   8; 11-bit with 4 */
  int32_t trav_order; /* persistent closest hit: 1 = batches of 2^18 rays or more are traced
                         grouped by origin (512 cells of the tree's root box, one counting
                         pass; results still at each ray's index), 0 = in input order */
  int32_t render_order; /* path segments after the camera's, their queue grouped by ray-origin
                           cell first: 0 never, 1 single-scene batches of 2^22 paths or more
                           (the default), 2 every single-scene batch; images are the same */
  int32_t pad;
} trace_tune;
int trace_get_tune(trace_ctx* ctx, trace_tune* out);
int trace_set_tune(trace_ctx* ctx, const trace_tune* tune);
/* Times closest hit on d_rays (device, n rays) for each workgroup size (64, 128, 256)
 * with one ray per lane, then with persistent lanes in input order, then with
 * persistent lanes and rays grouped by origin, keeps the fastest in the context, and
 * reports the per-candidate times in ms (9 entries, in that order). */
int trace_autotune_traversal(trace_ctx* ctx, const trace_bvh* bvh, const float* d_rays,
                             int64_t n, float* h_ms);
/* Times a render for each workgroup size, wavefront batch size (candidate batches
 * 2^22, 2^24, 2^26, 2^27 paths) and segment mode (fused: one segment per lane traced
 * and shaded; split: a persistent trace pass, then a shade pass) in turn and keeps the
 * fastest; h_ms receives 3 + 4 + 2 times. */
int trace_autotune_render(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh,
                          const trace_camera* cam, const trace_render_params* prm, float* h_ms);
/* Times the LBVH build of `mesh` (key_mode as trace_lbvh_build) for each key-sort
 * configuration (digit bits, keys per thread) = (8, 16), (8, 8), (8, 4), (10, 8),
 * (10, 4), (11, 4) -- 8, 8, 8, 7, 7 and 6 passes of 63-bit keys, tiles of 4096 to
 * 1024 keys -- keeps the fastest in the context and reports the times in ms (6
 * entries, in that order).  Every configuration builds the same tree. */
int trace_autotune_build(trace_ctx* ctx, const trace_mesh* mesh, int key_mode, float* h_ms);

#ifdef __cplusplus
}
#endif

#endif /* TRACE_HIP_H */
