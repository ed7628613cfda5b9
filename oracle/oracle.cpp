// ORACLE -- test infrastructure only.
//
// A CPU, FP64 restatement of the reference's hot path (ChSchmidt99/trace,
// pkg/pt).  It exists to check the HIP library: only tests/, the smoke check
// and bench.py's CPU-baseline leg load it, and never as the thing measured or
// shipped.  Every function cites the reference lines it restates.
//
// Deliberate, documented differences from the reference (all are places where
// the reference's own result is not defined by its source):
//   - the Morton sort is stable (ties keep triangle index order); the
//     reference's sort.Slice is unstable (lbvh.go:298), and equal keys land in
//     one leaf anyway (lbvh.go:134-145);
//   - a zero-extent axis quantises to 0 (Go's uint64(NaN) is platform defined);
//   - a one-triangle scene yields a root leaf (the reference dereferences a nil
//     parent, bvh.go:202-206);
//   - the closest-hit result carries the triangle index, which the reference's
//     hit record does not have (primitive.go:7-13).
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off), output oracle/_build/.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../trace_amd/csrc/philox.hpp"

namespace {

struct V3 {
  double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }  // math.go Dot
inline V3 cross(V3 a, V3 b) {                                                // math.go Cross
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// util.go:47-66
inline double min3(double a, double b, double c) {
  if (a <= b && a <= c) return a;
  if (b <= a && b <= c) return b;
  return c;
}
inline double max3(double a, double b, double c) {
  if (a >= b && a >= c) return a;
  if (b >= a && b >= c) return b;
  return c;
}

struct Box {
  double lo[3], hi[3];
};

// A primitive slot is a triangle (9 coordinates) or a Sphere: centre, radius,
// then NaN in the fifth value.
inline bool is_sphere(const double* p) { return p[4] != p[4]; }

// primitive.go:102-114 (box of a triangle); NewSphere's centre -/+ radius
// (primitive.go:31-40) for a sphere
inline Box tri_box(const double* p) {
  Box b;
  if (is_sphere(p)) {
    for (int a = 0; a < 3; ++a) {
      b.lo[a] = p[a] - p[3];
      b.hi[a] = p[a] + p[3];
    }
    return b;
  }
  for (int a = 0; a < 3; ++a) {
    b.lo[a] = min3(p[a], p[3 + a], p[6 + a]);
    b.hi[a] = max3(p[a], p[3 + a], p[6 + a]);
  }
  return b;
}

// aabb.add via MinVec/MaxVec (bounding.go:49-51, util.go:69-100)
inline Box box_union(const Box& a, const Box& b) {
  Box r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = a.lo[k] > b.lo[k] ? b.lo[k] : a.lo[k];
    r.hi[k] = a.hi[k] < b.hi[k] ? b.hi[k] : a.hi[k];
  }
  return r;
}

// bounding.go:45-47
inline double surface(const Box& b) {
  const double w = b.hi[0] - b.lo[0], h = b.hi[1] - b.lo[1], d = b.hi[2] - b.lo[2];
  return 2 * w * h + 2 * w * d + 2 * h * d;
}

// morton.go:12-20
inline uint64_t split_by3(uint64_t a) {
  uint64_t x = a & 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffULL;
  x = (x | x << 16) & 0x1f0000ff0000ffULL;
  x = (x | x << 8) & 0x100f00f00f00f00fULL;
  x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
  x = (x | x << 2) & 0x1249249249249249ULL;
  return x;
}

// lbvh.go:64-73 with the centroid of bounding.go:36-43
inline uint64_t morton(const double* p, const Box& enc) {
  const Box b = tri_box(p);
  uint64_t q[3];
  for (int a = 0; a < 3; ++a) {
    const double c = (b.lo[a] + b.hi[a]) * (1.0 / 2.0);
    const double delta = std::fabs(enc.lo[a] - c);
    const double extent = enc.hi[a] - enc.lo[a];
    const double f = delta / (extent / double(2097152 - 1));
    q[a] = (f >= 0.0) ? uint64_t(f) : 0;  // NaN (zero extent) -> 0
  }
  return split_by3(q[0]) | split_by3(q[1]) << 1 | split_by3(q[2]) << 2;  // morton.go:6-10
}

struct Node {
  bool leaf;
  int32_t first, last;  // sorted-position range (inclusive)
  int32_t child[2];
  Box box;
};

struct Tree {
  int64_t n = 0;
  std::vector<double> pos;
  std::vector<uint64_t> keys;  // sorted
  std::vector<int32_t> perm;   // sorted position -> triangle index
  std::vector<Node> nodes;
  int32_t root = -1;
};

// lbvh.go:184-203 findSplit (binary search for the first key with mask set)
int64_t find_split(const uint64_t* k, int64_t len, uint64_t mask) {
  if ((k[0] & mask) > 0) return 0;
  int64_t l = 0, r = len - 1;
  while (l <= r) {
    const int64_t m = (l + r) / 2;
    if ((k[m] & mask) == 0) {
      l = m + 1;
    } else if ((k[m] & mask) > 0 && (k[m - 1] & mask) > 0) {
      r = m - 1;
    } else {
      return m;
    }
  }
  return len;
}

// lbvh.go:133-177 lbvhJob.process (recursive form of the job queue)
int32_t build_range(Tree& t, int64_t b, int64_t e, uint64_t mask) {
  const uint64_t* k = t.keys.data();
  while (true) {
    if (k[b] == k[e - 1]) {  // isLeaf, lbvh.go:179-181
      Node nd{};
      nd.leaf = true;
      nd.first = int32_t(b);
      nd.last = int32_t(e - 1);
      nd.child[0] = nd.child[1] = -1;
      t.nodes.push_back(nd);
      return int32_t(t.nodes.size() - 1);
    }
    const int64_t s = find_split(k + b, e - b, mask);
    if (s == 0 || s == e - b) {
      mask >>= 1;  // no node for an empty split (lbvh.go:150-159)
      continue;
    }
    Node nd{};
    nd.leaf = false;
    nd.first = int32_t(b);
    nd.last = int32_t(e - 1);
    t.nodes.push_back(nd);
    const int32_t id = int32_t(t.nodes.size() - 1);
    const int32_t c0 = build_range(t, b, b + s, mask >> 1);
    const int32_t c1 = build_range(t, b + s, e, mask >> 1);
    t.nodes[id].child[0] = c0;
    t.nodes[id].child[1] = c1;
    return id;
  }
}

// bvh.go:201-222 (values only; the union is exact so order does not matter)
Box refit(Tree& t, int32_t id) {
  Node& nd = t.nodes[id];
  if (nd.leaf) {
    Box b = tri_box(&t.pos[9 * size_t(t.perm[nd.first])]);
    for (int32_t i = nd.first + 1; i <= nd.last; ++i)
      b = box_union(b, tri_box(&t.pos[9 * size_t(t.perm[i])]));  // enclosingSlice
    nd.box = b;
    return b;
  }
  const Box b0 = refit(t, nd.child[0]);
  const Box b1 = refit(t, nd.child[1]);
  t.nodes[id].box = box_union(b0, b1);
  return t.nodes[id].box;
}

// bvh.go:226-239
double cost_of(const Tree& t, int32_t id) {
  const Node& nd = t.nodes[id];
  if (nd.leaf) return 1.0 * double(nd.last - nd.first + 1);
  double acc = 0.0;
  const double sa = surface(nd.box);
  for (int c = 0; c < 2; ++c) {
    if (sa != 0) acc += surface(t.nodes[nd.child[c]].box) / sa * cost_of(t, nd.child[c]);
  }
  return 2.0 + acc;
}

// ---- traversal ------------------------------------------------------------
struct Ray {
  V3 o, d, inv;
  int sign[3];
};

// camera.go:113-135 newRay
Ray make_ray(const float* r) {
  Ray ray;
  ray.o = {r[0], r[1], r[2]};
  ray.d = {r[4], r[5], r[6]};
  ray.inv = {1.0 / ray.d.x, 1.0 / ray.d.y, 1.0 / ray.d.z};
  ray.sign[0] = ray.inv.x < 0;
  ray.sign[1] = ray.inv.y < 0;
  ray.sign[2] = ray.inv.z < 0;
  return ray;
}

// bounding.go:53-91 (Williams et al. slab test, NaN never rejects)
bool box_hit(const Box& b, const Ray& r, double tmin, double tmax) {
  const double* bx[2] = {b.lo, b.hi};
  double txmin = (bx[r.sign[0]][0] - r.o.x) * r.inv.x;
  double txmax = (bx[1 - r.sign[0]][0] - r.o.x) * r.inv.x;
  const double tymin = (bx[r.sign[1]][1] - r.o.y) * r.inv.y;
  const double tymax = (bx[1 - r.sign[1]][1] - r.o.y) * r.inv.y;
  if (txmin > tymax || tymin > txmax) return false;
  if (tymin > txmin) txmin = tymin;
  if (tymax < txmax) txmax = tymax;
  const double tzmin = (bx[r.sign[2]][2] - r.o.z) * r.inv.z;
  const double tzmax = (bx[1 - r.sign[2]][2] - r.o.z) * r.inv.z;
  if (txmin > tzmax || tzmin > txmax) return false;
  if (tzmin > txmin) txmin = tzmin;
  if (tzmax < txmax) txmax = tzmax;
  if (txmax < tmin || txmin > tmax) return false;
  return true;
}

struct HitRec {
  double t;
  int32_t prim;  // primitive index
  double u, v;
  bool front;
  bool sphere = false;  // a Sphere hit: n holds its facing normal
  V3 n{0, 0, 0};
};

// Sphere.intersected (primitive.go:56-86): the nearer root strictly inside
// (tMin, tMax), else the farther; the normal (point - centre) * (1 / radius),
// flipped to face the ray.
bool sphere_hit(const double* p, const Ray& r, double tmin, double tmax, HitRec& h, int32_t prim) {
  const V3 c{p[0], p[1], p[2]};
  const double rad = p[3];
  const V3 oc = sub(r.o, c);
  const double dn = std::sqrt(dot(r.d, r.d));
  const double a = dn * dn;
  const double hb = dot(oc, r.d);
  const double on = std::sqrt(dot(oc, oc));
  const double cc = on * on - rad * rad;
  const double disc = hb * hb - a * cc;
  if (disc < 0) return false;
  const double sq = std::sqrt(disc);
  double t = (-hb - sq) / a;
  if (t <= tmin || t >= tmax) {
    t = (-hb + sq) / a;
    if (t <= tmin || t >= tmax) return false;
  }
  const V3 pt{r.o.x + r.d.x * t, r.o.y + r.d.y * t, r.o.z + r.d.z * t};
  const double inv = 1 / rad;
  V3 n{(pt.x - c.x) * inv, (pt.y - c.y) * inv, (pt.z - c.z) * inv};
  const bool front = dot(r.d, n) < 0;
  if (!front) n = {-n.x, -n.y, -n.z};
  h.t = t;
  h.prim = prim;
  h.u = h.v = 0;
  h.front = front;
  h.sphere = true;
  h.n = n;
  return true;
}

// primitive.go:172-208 Moeller-Trumbore, util.go:10-12 ApproxZero
bool tri_hit(const double* p, const Ray& r, double tmin, double tmax, HitRec& h, int32_t prim) {
  if (is_sphere(p)) return sphere_hit(p, r, tmin, tmax, h, prim);
  const V3 v0{p[0], p[1], p[2]};
  const V3 e1 = sub(V3{p[3], p[4], p[5]}, v0);
  const V3 e2 = sub(V3{p[6], p[7], p[8]}, v0);
  const V3 pv = cross(r.d, e2);
  const double det = dot(e1, pv);
  if (std::fabs(det) <= 1e-7) return false;
  const double inv = 1 / det;
  const V3 tv = sub(r.o, v0);
  const double u = dot(tv, pv) * inv;
  if (u < 0 || u > 1) return false;
  const V3 q = cross(tv, e1);
  const double v = dot(r.d, q) * inv;
  if (v < 0 || u + v > 1) return false;
  const double t = dot(e2, q) * inv;
  if (t < tmin || t > tmax) return false;
  h.t = t;
  h.prim = prim;
  h.u = u;
  h.v = v;
  h.front = det > 0;
  h.sphere = false;
  return true;
}

// bvh.go:179-199 (recursive, children in stored order; each accepted hit
// tightens tMax for the siblings that follow)
bool node_hit(const Tree& t, int32_t id, const Ray& r, double tmin, double tmax, HitRec& h) {
  const Node& nd = t.nodes[id];
  if (!box_hit(nd.box, r, tmin, tmax)) return false;
  bool did = false;
  if (nd.leaf) {
    for (int32_t i = nd.first; i <= nd.last; ++i) {
      const int32_t prim = t.perm[i];
      if (tri_hit(&t.pos[9 * size_t(prim)], r, tmin, h.t, h, prim)) did = true;
    }
  } else {
    for (int c = 0; c < 2; ++c)
      if (node_hit(t, nd.child[c], r, tmin, h.t, h)) did = true;
  }
  return did;
}

template <typename F>
void parallel_for(int64_t n, int threads, F f) {
  if (threads <= 1 || n < 1024) {
    f(int64_t(0), n);
    return;
  }
  std::vector<std::thread> ts;
  const int64_t chunk = (n + threads - 1) / threads;
  for (int i = 0; i < threads; ++i) {
    const int64_t b = i * chunk, e = std::min<int64_t>(n, b + chunk);
    if (b >= e) break;
    ts.emplace_back([=] { f(b, e); });
  }
  for (auto& th : ts) th.join();
}

}  // namespace

extern "C" {

typedef struct Tree oracle_tree;

uint64_t oracle_split_by3(uint64_t a) { return split_by3(a); }

// bounding.go:19-25
void oracle_enclosing(const double* pos, int64_t n, double* out6) {
  Box b = tri_box(pos);
  for (int64_t i = 1; i < n; ++i) b = box_union(b, tri_box(pos + 9 * i));
  for (int a = 0; a < 3; ++a) {
    out6[a] = b.lo[a];
    out6[3 + a] = b.hi[a];
  }
}

// lbvh.go:33-73: code per triangle; key_mode 1 keeps code >> 33 (30 bits)
void oracle_morton(const double* pos, int64_t n, const double* qbox6, int key_mode,
                   uint64_t* keys) {
  Box enc;
  for (int a = 0; a < 3; ++a) {
    enc.lo[a] = qbox6[a];
    enc.hi[a] = qbox6[3 + a];
  }
  for (int64_t i = 0; i < n; ++i) {
    uint64_t c = morton(pos + 9 * i, enc);
    keys[i] = key_mode == 1 ? (c >> 33) : c;
  }
}

// lbvh.go:216-301 result: pairs ascending by code (stable)
void oracle_sort(const uint64_t* keys, int64_t n, uint64_t* sorted, int32_t* perm) {
  std::vector<int32_t> idx(n);
  for (int64_t i = 0; i < n; ++i) idx[i] = int32_t(i);
  std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return keys[a] < keys[b]; });
  for (int64_t i = 0; i < n; ++i) {
    perm[i] = idx[i];
    sorted[i] = keys[idx[i]];
  }
}

// LBVH (lbvh.go:17-25): Morton codes, sort, top-down split from bit 62, refit.
oracle_tree* oracle_lbvh_build(const double* pos, int64_t n, const double* enclosing6,
                               int key_mode) {
  if (n <= 0) return nullptr;
  Tree* t = new Tree();
  t->n = n;
  t->pos.assign(pos, pos + 9 * n);
  double box[6];
  if (enclosing6) std::memcpy(box, enclosing6, sizeof(box));
  else oracle_enclosing(pos, n, box);
  std::vector<uint64_t> codes(n);
  oracle_morton(pos, n, box, key_mode, codes.data());
  t->keys.resize(n);
  t->perm.resize(n);
  oracle_sort(codes.data(), n, t->keys.data(), t->perm.data());
  t->nodes.reserve(2 * n);
  t->root = build_range(*t, 0, n, uint64_t(1) << 62);
  refit(*t, t->root);
  return t;
}

void oracle_tree_free(oracle_tree* t) { delete t; }

int64_t oracle_tree_node_count(const oracle_tree* t) { return int64_t(t->nodes.size()); }

void oracle_tree_sorted(const oracle_tree* t, uint64_t* keys, int32_t* perm) {
  std::memcpy(keys, t->keys.data(), t->n * sizeof(uint64_t));
  std::memcpy(perm, t->perm.data(), t->n * sizeof(int32_t));
}

// Preorder, children[0] first: rec = kind, first, last, n_children; box FP64.
void oracle_tree_canonical(const oracle_tree* t, int32_t* rec, double* box) {
  std::vector<int32_t> st{t->root};
  int64_t k = 0;
  while (!st.empty()) {
    const int32_t id = st.back();
    st.pop_back();
    const Node& nd = t->nodes[id];
    rec[4 * k + 0] = nd.leaf ? 1 : 0;
    rec[4 * k + 1] = nd.first;
    rec[4 * k + 2] = nd.last;
    rec[4 * k + 3] = nd.leaf ? 0 : 2;
    for (int a = 0; a < 3; ++a) {
      box[6 * k + a] = nd.box.lo[a];
      box[6 * k + 3 + a] = nd.box.hi[a];
    }
    if (!nd.leaf) {
      st.push_back(nd.child[1]);
      st.push_back(nd.child[0]);
    }
    ++k;
  }
}

double oracle_tree_cost(const oracle_tree* t) { return cost_of(*t, t->root); }

// BVH.intersected (bvh.go:74-77) for a batch of rays (8 floats each); t is FP64.
void oracle_closest_hit(const oracle_tree* t, const float* rays, int64_t nrays, int32_t* hit,
                        double* tout, int threads) {
  parallel_for(nrays, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const float* rr = rays + 8 * i;
      const Ray r = make_ray(rr);
      HitRec h{};
      h.t = rr[7];
      h.prim = -1;
      node_hit(*t, t->root, r, rr[3], rr[7], h);
      hit[i] = h.prim;
      tout[i] = h.t;
    }
  });
}

// All triangles in index order, no tree (pins the traversal on small inputs).
void oracle_closest_hit_brute(const double* pos, int64_t n, const float* rays, int64_t nrays,
                              int32_t* hit, double* tout, int threads) {
  parallel_for(nrays, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const float* rr = rays + 8 * i;
      const Ray r = make_ray(rr);
      HitRec h{};
      h.t = rr[7];
      h.prim = -1;
      for (int64_t k = 0; k < n; ++k) tri_hit(pos + 9 * k, r, rr[3], h.t, h, int32_t(k));
      hit[i] = h.prim;
      tout[i] = h.t;
    }
  });
}

void oracle_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                   uint32_t* out4) {
  const Philox4 r = philox4x32_10(c0, c1, c2, c3, k0, k1);
  std::memcpy(out4, r.v, sizeof(r.v));
}

}  // extern "C"

// ---- PHR (phr.go) -----------------------------------------------------------
// Progressive Hierarchical Refinement over the LBVH.  Deterministic: the
// initial cut is taken in FIFO order from the root (what the reference's
// channel queue does with one worker while it holds < 1024 nodes,
// phr.go:257-311), and subtrees are built depth first.  Every other step is
// order independent in the reference.
namespace {

struct PNode {
  int kind = 0;               // 0 inner (PHR branch), 1 PHR-made leaf, 2 reused LBVH node
  int32_t aux = -1;           // kind 2: LBVH node index
  std::vector<int32_t> kids;  // kind 0
  std::vector<int32_t> prims; // kind 1: sorted positions
  Box box;
};

struct Ref {  // a member of a cut: an LBVH node or a PHR-made leaf
  int32_t id;
  bool made;
};

struct Cut {
  std::vector<Ref> nodes;
  Box bounding;
  int depth = 0;
};

struct Phr {
  const Tree* aux;
  double alpha, delta, S;
  int bf;
  std::vector<PNode> out;
  std::vector<PNode> made;      // PHR-made leaves referenced from cuts
  std::vector<int64_t> size;    // LBVH subtree sizes (bvh.go:159-167)
  int64_t split_work = 0;       // BuildWithCost's counter (phr.go:221)
  int64_t max_split = 0;        // largest cut handed to SweepSAH
  // Optional record of every SweepSAH call, for tests: per call the ints
  // m, split, nl, nr, the number of cuts the job holds and the index of the
  // one split, each cut's size, then (first, last) sorted-position range of each member
  // (-1, -1 for a PHR-made leaf), the members taken left and right (indices
  // into the member list, in the order the split cuts hold them); the doubles
  // hold the cut's bounding box, then each member's box.
  bool tracing = false;
  std::vector<int32_t> trace_i;
  std::vector<double> trace_d;

  const Box& box_of(Ref r) const { return r.made ? made[r.id].box : aux->nodes[r.id].box; }
  bool is_leaf(Ref r) const { return r.made || aux->nodes[r.id].leaf; }
  int64_t subtree(Ref r) const { return r.made ? 1 : size[r.id]; }

  // phr.go:356-358
  double threshold(int depth) const { return S / std::pow(2.0, alpha * double(depth) + delta); }

  int64_t compute_size(int32_t id) {
    const Node& nd = aux->nodes[id];
    int64_t s = 1;
    if (!nd.leaf) s += compute_size(nd.child[0]) + compute_size(nd.child[1]);
    size[id] = s;
    return s;
  }

  // phr.go:340-352: leaves of each node in order, their prims concatenated
  Ref make_leaf(const Box& bounding, const std::vector<Ref>& nodes) {
    PNode leaf;
    leaf.kind = 1;
    leaf.box = bounding;
    for (const Ref& r : nodes) {
      if (r.made) {
        const auto& p = made[r.id].prims;
        leaf.prims.insert(leaf.prims.end(), p.begin(), p.end());
        continue;
      }
      std::vector<int32_t> st{r.id};
      while (!st.empty()) {
        const int32_t id = st.back();
        st.pop_back();
        const Node& nd = aux->nodes[id];
        if (nd.leaf) {
          for (int32_t i = nd.first; i <= nd.last; ++i) leaf.prims.push_back(i);
        } else {
          st.push_back(nd.child[1]);
          st.push_back(nd.child[0]);
        }
      }
    }
    made.push_back(leaf);
    return Ref{int32_t(made.size() - 1), true};
  }

  // phr.go:313-338
  Cut refined(const Cut& cut) {
    Cut r;
    r.bounding = cut.bounding;
    r.depth = cut.depth + 1;
    const double th = threshold(cut.depth);
    for (const Ref& n : cut.nodes) {
      if (is_leaf(n) || surface(box_of(n)) < th) {
        r.nodes.push_back(n);
      } else {
        const Node& nd = aux->nodes[n.id];
        r.nodes.push_back(Ref{nd.child[0], false});
        r.nodes.push_back(Ref{nd.child[1], false});
      }
    }
    if (r.nodes.size() == 1) r.nodes[0] = make_leaf(cut.bounding, r.nodes);
    return r;
  }

  struct Sah {
    double cost;
    int64_t split;  // 0 = no split (left = all)
    Box lbox, rbox;
  };

  // phr.go:414-459
  Sah min_cost(const std::vector<Ref>& s, const Box& bounding) const {
    const int64_t m = int64_t(s.size());
    std::vector<double> right_cost(m, 0.0);
    std::vector<Box> right_box(m);
    Box sr = box_of(s[m - 1]);
    int64_t count = 0;
    for (int64_t i = m - 1; i > 0; --i) {
      sr = box_union(sr, box_of(s[i]));
      count += subtree(s[i]);
      right_cost[i] = surface(sr) * double(count);
      right_box[i] = sr;
    }
    count += subtree(s[0]);
    Sah best{surface(bounding) * double(count), 0, bounding, Box{}};
    count = subtree(s[0]);
    Box sl = box_of(s[0]);
    for (int64_t i = 1; i < m; ++i) {
      const double c = right_cost[i] + surface(sl) * double(count);
      if (c < best.cost) best = Sah{c, i, sl, right_box[i]};
      sl = box_union(sl, box_of(s[i]));
      count += subtree(s[i]);
    }
    return best;
  }

  std::vector<Ref> sorted_by(std::vector<Ref> v, int axis) const {
    std::stable_sort(v.begin(), v.end(), [&](const Ref& a, const Ref& b) {
      const Box& ba = box_of(a);
      const Box& bb = box_of(b);
      return (ba.lo[axis] + ba.hi[axis]) * (1.0 / 2.0) < (bb.lo[axis] + bb.hi[axis]) * (1.0 / 2.0);
    });
    return v;
  }

  // phr.go:368-405: returns false when no split beats keeping the cut whole
  bool sweep_sah(const Cut& cut, Cut& left, Cut& right) {
    const std::vector<Ref> xs = sorted_by(cut.nodes, 0);
    const std::vector<Ref> ys = sorted_by(xs, 1);
    const Sah x = min_cost(xs, cut.bounding);
    const Sah y = min_cost(ys, cut.bounding);
    const std::vector<Ref>* chosen;
    Sah best;
    if (x.cost < y.cost) {
      const std::vector<Ref> zs = sorted_by(ys, 2);
      const Sah z = min_cost(zs, cut.bounding);
      if (x.cost < z.cost) { best = x; chosen = &xs; finish(cut, *chosen, best, left, right); }
      else { best = z; finish(cut, zs, best, left, right); }
    } else {
      const std::vector<Ref> zs = sorted_by(xs, 2);
      const Sah z = min_cost(zs, cut.bounding);
      if (y.cost < z.cost) { best = y; chosen = &ys; finish(cut, *chosen, best, left, right); }
      else { best = z; finish(cut, zs, best, left, right); }
    }
    return best.split != 0;
  }

  void finish(const Cut& cut, const std::vector<Ref>& s, const Sah& b, Cut& left, Cut& right) {
    left.depth = right.depth = cut.depth;
    if (b.split == 0) {
      left.nodes = s;
      left.bounding = cut.bounding;
      return;
    }
    left.nodes.assign(s.begin(), s.begin() + b.split);
    left.bounding = b.lbox;
    right.nodes.assign(s.begin() + b.split, s.end());
    right.bounding = b.rbox;
  }

  void record(const std::vector<Cut>& cuts, size_t mi, bool split, const Cut& l, const Cut& r) {
    const Cut& cut = cuts[mi];
    const int32_t m = int32_t(cut.nodes.size());
    auto index_of = [&](const Ref& x) {
      for (int32_t k = 0; k < m; ++k)
        if (cut.nodes[k].id == x.id && cut.nodes[k].made == x.made) return k;
      return int32_t(-1);
    };
    trace_i.push_back(m);
    trace_i.push_back(split ? 1 : 0);
    trace_i.push_back(split ? int32_t(l.nodes.size()) : 0);
    trace_i.push_back(split ? int32_t(r.nodes.size()) : 0);
    trace_i.push_back(int32_t(cuts.size()));
    trace_i.push_back(int32_t(mi));
    for (const Cut& c : cuts) trace_i.push_back(int32_t(c.nodes.size()));
    for (const Ref& x : cut.nodes) {
      trace_i.push_back(x.made ? -1 : aux->nodes[x.id].first);
      trace_i.push_back(x.made ? -1 : aux->nodes[x.id].last);
    }
    if (split) {
      for (const Ref& x : l.nodes) trace_i.push_back(index_of(x));
      for (const Ref& x : r.nodes) trace_i.push_back(index_of(x));
    }
    for (int a = 0; a < 3; ++a) trace_d.push_back(cut.bounding.lo[a]);
    for (int a = 0; a < 3; ++a) trace_d.push_back(cut.bounding.hi[a]);
    for (const Ref& x : cut.nodes) {
      const Box& b = box_of(x);
      for (int a = 0; a < 3; ++a) trace_d.push_back(b.lo[a]);
      for (int a = 0; a < 3; ++a) trace_d.push_back(b.hi[a]);
    }
  }

  int32_t attach(Ref r) {
    PNode p;
    if (r.made) {
      p = made[r.id];
    } else {
      p.kind = 2;
      p.aux = r.id;
      p.box = aux->nodes[r.id].box;
    }
    out.push_back(p);
    return int32_t(out.size() - 1);
  }

  // phr.go:92-152 (and the split counter of buildSubTreeCost, phr.go:221)
  int32_t build(const Cut& cut) {
    if (cut.nodes.size() <= 1) return attach(cut.nodes[0]);
    std::vector<Cut> cuts{cut};
    while (int(cuts.size()) < bf) {
      size_t mx = 0, mi = 0;
      for (size_t i = 0; i < cuts.size(); ++i)
        if (cuts[i].nodes.size() > mx) { mx = cuts[i].nodes.size(); mi = i; }
      if (mx <= 1) break;
      split_work += int64_t(cuts[mi].nodes.size());
      max_split = std::max<int64_t>(max_split, int64_t(cuts[mi].nodes.size()));
      Cut l, r;
      const bool split = sweep_sah(cuts[mi], l, r);
      if (tracing) record(cuts, mi, split, l, r);
      if (split) {
        Cut rl = refined(l), rr = refined(r);
        cuts[mi] = rl;
        cuts.push_back(rr);
      } else {
        Cut c;
        c.nodes.push_back(make_leaf(l.bounding, l.nodes));
        cuts[mi] = c;
      }
    }
    PNode br;
    br.kind = 0;
    br.box = cut.bounding;
    out.push_back(br);
    const int32_t id = int32_t(out.size() - 1);
    for (const Cut& c : cuts) {
      const int32_t k = build(c);
      out[id].kids.push_back(k);
    }
    return id;
  }

  // phr.go:257-311 in FIFO order
  Cut initial_cut(int64_t max_cut) {
    Cut cut;
    cut.bounding = aux->nodes[aux->root].box;
    cut.depth = 1;
    const double th = threshold(0);
    int64_t counter = 0;
    std::vector<int32_t> fifo{aux->root};
    for (size_t h = 0; h < fifo.size(); ++h) {
      const int32_t id = fifo[h];
      const Node& nd = aux->nodes[id];
      if (nd.leaf) {
        cut.nodes.push_back(Ref{id, false});
      } else if (surface(nd.box) > th) {
        counter += 1;  // len(children) - 1
        if (counter >= max_cut) {
          cut.nodes.push_back(Ref{id, false});
        } else {
          fifo.push_back(nd.child[0]);
          fifo.push_back(nd.child[1]);
        }
      } else {
        cut.nodes.push_back(Ref{id, false});
      }
    }
    return cut;
  }
};

struct PhrTree {
  const Tree* aux;
  std::vector<PNode> nodes;
  int32_t root;
  int64_t split_work;
  int64_t initial_cut;
  int64_t max_split;
  std::vector<int32_t> trace_i;
  std::vector<double> trace_d;
  std::vector<int32_t> cut_ranges;  // the initial cut's members: (first, last) each
};

double phr_cost(const PhrTree& t, const Tree& aux, int32_t id);

double aux_cost(const Tree& aux, int32_t id) { return cost_of(aux, id); }

double phr_cost(const PhrTree& t, const Tree& aux, int32_t id) {
  const PNode& p = t.nodes[id];
  if (p.kind == 2) return aux_cost(aux, p.aux);
  if (p.kind == 1) return 1.0 * double(p.prims.size());
  double acc = 0.0;
  const double sa = surface(p.box);
  for (int32_t k : p.kids) {
    const PNode& c = t.nodes[k];
    const Box& cb = c.kind == 2 ? aux.nodes[c.aux].box : c.box;
    if (sa != 0) acc += surface(cb) / sa * phr_cost(t, aux, k);
  }
  return 2.0 + acc;
}

}  // namespace

extern "C" {

typedef struct PhrTree oracle_phr;

// PhrBuilder.BuildFromAuxilary (phr.go:47-84) on the oracle's LBVH.
static oracle_phr* phr_build(const oracle_tree* aux, double alpha, double delta, int branching, int64_t max_cut,
                             bool tracing) {
  Phr b;
  b.tracing = tracing;
  b.aux = aux;
  b.alpha = alpha;
  b.delta = delta;
  b.bf = branching;
  b.S = surface(aux->nodes[aux->root].box);
  b.size.assign(aux->nodes.size(), 0);
  b.compute_size(aux->root);
  Cut cut = b.initial_cut(max_cut);
  const int64_t n_cut = int64_t(cut.nodes.size());
  b.split_work = n_cut;  // BuildWithCost counts the initial cut (phr.go:160)
  // The temporary parent of phr.go:64-65 takes job.cut.bounding = root box.
  std::vector<int32_t> ranges;
  for (const Ref& x : cut.nodes) {
    ranges.push_back(aux->nodes[x.id].first);
    ranges.push_back(aux->nodes[x.id].last);
  }
  const int32_t root = b.build(cut);
  PhrTree* t = new PhrTree{aux,       std::move(b.out),     root, b.split_work, n_cut, b.max_split,
                           std::move(b.trace_i), std::move(b.trace_d), std::move(ranges)};
  return t;
}

oracle_phr* oracle_phr_build(const oracle_tree* aux, double alpha, double delta, int branching,
                             int64_t max_cut) {
  return phr_build(aux, alpha, delta, branching, max_cut, false);
}

// oracle_phr_build that also records every SweepSAH call (Phr::record).
oracle_phr* oracle_phr_build_traced(const oracle_tree* aux, double alpha, double delta, int branching,
                                    int64_t max_cut) {
  return phr_build(aux, alpha, delta, branching, max_cut, true);
}

// Copies the SweepSAH record (null outputs: sizes only).
void oracle_phr_trace(const oracle_phr* t, int32_t* ints, double* dbls, int64_t* n_ints, int64_t* n_dbls) {
  *n_ints = int64_t(t->trace_i.size());
  *n_dbls = int64_t(t->trace_d.size());
  if (ints) std::copy(t->trace_i.begin(), t->trace_i.end(), ints);
  if (dbls) std::copy(t->trace_d.begin(), t->trace_d.end(), dbls);
}

// The initial cut in its order: (first, last) sorted-position range per member.
int64_t oracle_phr_initial_ranges(const oracle_phr* t, int32_t* out) {
  if (out) std::copy(t->cut_ranges.begin(), t->cut_ranges.end(), out);
  return int64_t(t->cut_ranges.size() / 2);
}

void oracle_phr_free(oracle_phr* t) { delete t; }

double oracle_phr_cost(const oracle_phr* t) { return phr_cost(*t, *t->aux, t->root); }

}  // extern "C"

namespace {
// bvh.go:179-199 over the PHR tree: node box first, then children in stored
// order (reused LBVH subtrees walk the LBVH), leaf triangles in list order.
bool phr_node_hit(const PhrTree& t, int32_t id, const Ray& r, double tmin, double tmax, HitRec& h) {
  const PNode& p = t.nodes[id];
  if (p.kind == 2) return node_hit(*t.aux, p.aux, r, tmin, tmax, h);
  if (!box_hit(p.box, r, tmin, tmax)) return false;
  bool did = false;
  if (p.kind == 1) {
    for (int32_t pos : p.prims) {
      const int32_t prim = t.aux->perm[pos];
      if (tri_hit(&t.aux->pos[9 * size_t(prim)], r, tmin, h.t, h, prim)) did = true;
    }
  } else {
    for (int32_t k : p.kids)
      if (phr_node_hit(t, k, r, tmin, h.t, h)) did = true;
  }
  return did;
}
}  // namespace

extern "C" {

// BVH.intersected on the PHR tree for a batch of rays.
void oracle_phr_closest_hit(const oracle_phr* t, const float* rays, int64_t nrays, int32_t* hit,
                            double* tout, int threads) {
  parallel_for(nrays, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const float* rr = rays + 8 * i;
      const Ray r = make_ray(rr);
      HitRec h{};
      h.t = rr[7];
      h.prim = -1;
      phr_node_hit(*t, t->root, r, rr[3], rr[7], h);
      hit[i] = h.prim;
      tout[i] = h.t;
    }
  });
}

}  // extern "C"

namespace {
// camera.go:113-135 newRay from FP64 components.
Ray make_ray64(const double* r) {
  Ray ray;
  ray.o = {r[0], r[1], r[2]};
  ray.d = {r[4], r[5], r[6]};
  ray.inv = {1.0 / ray.d.x, 1.0 / ray.d.y, 1.0 / ray.d.z};
  ray.sign[0] = ray.inv.x < 0;
  ray.sign[1] = ray.inv.y < 0;
  ray.sign[2] = ray.inv.z < 0;
  return ray;
}

// bvh.go:21-44 traversalSteps: an explicit stack that pushes a node's children
// in stored order (so the last child is popped first); an inner node whose box
// the ray enters under the current best t counts one step and pushes its
// children, a leaf tests its triangles, which tighten t.  A stack entry is an
// LBVH node (aux = true) or a PHR node; PHR nodes of kind 2 are LBVH nodes.
struct StepRef {
  int32_t id;
  bool aux;
};

int32_t steps_walk(const Tree& lb, const PhrTree* ph, StepRef root, const Ray& r, double tmin, double tmax) {
  std::vector<StepRef> st{root};
  int32_t count = 0;
  HitRec h{};
  h.t = tmax;
  h.prim = -1;
  while (!st.empty()) {
    StepRef e = st.back();
    st.pop_back();
    if (!e.aux && ph->nodes[e.id].kind == 2) e = StepRef{ph->nodes[e.id].aux, true};
    if (e.aux) {
      const Node& nd = lb.nodes[e.id];
      if (!box_hit(nd.box, r, tmin, h.t)) continue;
      if (nd.leaf) {
        for (int32_t i = nd.first; i <= nd.last; ++i) {
          const int32_t prim = lb.perm[i];
          tri_hit(&lb.pos[9 * size_t(prim)], r, tmin, h.t, h, prim);
        }
      } else {
        st.push_back(StepRef{nd.child[0], true});
        st.push_back(StepRef{nd.child[1], true});
        ++count;
      }
    } else {
      const PNode& p = ph->nodes[e.id];
      if (!box_hit(p.box, r, tmin, h.t)) continue;
      if (p.kind == 1) {
        for (int32_t pos : p.prims) {
          const int32_t prim = lb.perm[pos];
          tri_hit(&lb.pos[9 * size_t(prim)], r, tmin, h.t, h, prim);
        }
      } else {
        for (int32_t k : p.kids) st.push_back(StepRef{k, false});
        ++count;
      }
    }
  }
  return count;
}

// render.go:250-258 DefaultTraversalCountShader
void count_colour(int32_t count, int32_t threshold, float* rgb) {
  if (count > threshold) {
    rgb[0] = float(double(count) / double(threshold) * 2);
    rgb[1] = 0.0f;
    rgb[2] = 0.0f;
    return;
  }
  const double f = double(count) / (double(threshold) * 1.25);
  rgb[0] = 0.0f;
  rgb[1] = float(f);
  rgb[2] = float(1 - f);
}

// HeatMapRenderer.RenderToBuffer (render.go:167-197): pixel (x, y) casts the
// camera ray through (x / (W-1), y / (H-1)) (camera.go:95-97) with tMin 0.001
// and tMax = math.MaxFloat64; pixel index y * W + x.
void heatmap(const Tree& lb, const PhrTree* ph, StepRef root, const double* cam, int32_t W, int32_t H,
             int32_t threshold, float* rgb, int32_t* counts, int threads) {
  parallel_for(int64_t(W) * H, threads, [&](int64_t b, int64_t e) {
    for (int64_t pix = b; pix < e; ++pix) {
      const int32_t x = int32_t(pix % W), y = int32_t(pix / W);
      const double u = double(x) / double(W - 1);
      const double v = double(y) / double(H - 1);
      const double r8[8] = {cam[0], cam[1], cam[2], 0.0,
                            cam[3] + cam[6] * u + cam[9] * v - cam[0],
                            cam[4] + cam[7] * u + cam[10] * v - cam[1],
                            cam[5] + cam[8] * u + cam[11] * v - cam[2], 0.0};
      const int32_t c = steps_walk(lb, ph, root, make_ray64(r8), 0.001, DBL_MAX);
      if (counts) counts[pix] = c;
      count_colour(c, threshold, rgb + 3 * pix);
    }
  });
}
}  // namespace

extern "C" {

// BVH.traversalSteps (bvh.go:21-44) for a batch of FP64 rays (o, tMin, d, tMax).
void oracle_traversal_steps(const oracle_tree* t, const double* rays, int64_t nrays, int32_t* counts,
                            int threads) {
  parallel_for(nrays, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const double* rr = rays + 8 * i;
      counts[i] = steps_walk(*t, nullptr, StepRef{t->root, true}, make_ray64(rr), rr[3], rr[7]);
    }
  });
}

void oracle_phr_traversal_steps(const oracle_phr* t, const double* rays, int64_t nrays, int32_t* counts,
                                int threads) {
  parallel_for(nrays, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const double* rr = rays + 8 * i;
      counts[i] = steps_walk(*t->aux, t, StepRef{t->root, false}, make_ray64(rr), rr[3], rr[7]);
    }
  });
}

// HeatMapRenderer.RenderToBuffer into RGB floats (one sample per pixel, so the
// PixelBuffer mean is the colour) and the per-pixel counts (optional).
void oracle_heatmap(const oracle_tree* t, const double* cam, int32_t W, int32_t H, int32_t threshold,
                    float* rgb, int32_t* counts, int threads) {
  heatmap(*t, nullptr, StepRef{t->root, true}, cam, W, H, threshold, rgb, counts, threads);
}

void oracle_phr_heatmap(const oracle_phr* t, const double* cam, int32_t W, int32_t H, int32_t threshold,
                        float* rgb, int32_t* counts, int threads) {
  heatmap(*t->aux, t, StepRef{t->root, false}, cam, W, H, threshold, rgb, counts, threads);
}

}  // extern "C"

namespace {
// ---- scene compile (scene.go:95-151, primitive.go:152-170, math.go:144-320) ----
struct M4 {
  double m[16];
};

M4 m4_identity() {
  M4 r{};
  r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1;
  return r;
}

// MultiplyMatrix (math.go:299-320)
M4 m4_mul(const M4& A, const M4& B) {
  const double* a = A.m;
  const double* b = B.m;
  M4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      r.m[4 * i + j] = a[4 * i + 0] * b[0 + j] + a[4 * i + 1] * b[4 + j] + a[4 * i + 2] * b[8 + j] +
                       a[4 * i + 3] * b[12 + j];
  return r;
}

M4 m4_transpose(const M4& A) {
  M4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r.m[4 * j + i] = A.m[4 * i + j];
  return r;
}

// Inverse (math.go:164-297): adjugate by cofactors, each a signed sum of six
// triple products evaluated left to right, then scaled by 1 / det.
M4 m4_inverse(const M4& A) {
  const double* m = A.m;
  double v[16];
  v[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
         m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  v[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
         m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  v[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
         m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  v[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
          m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  v[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
         m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  v[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
         m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  v[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
         m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  v[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
          m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  v[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
         m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  v[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
         m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  v[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
          m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  v[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
          m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  v[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
         m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  v[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
         m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  v[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
          m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  v[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
          m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  double det = m[0] * v[0] + m[1] * v[4] + m[2] * v[8] + m[3] * v[12];
  det = 1.0 / det;
  M4 r;
  for (int k = 0; k < 16; ++k) r.m[k] = v[k] * det;
  return r;
}

// Vector4.Transformed of the point (x, y, z, 1), then ToV3 (math.go:108-162)
void m4_point(const M4& A, const double* p, double* out) {
  const double* m = A.m;
  for (int r = 0; r < 3; ++r) out[r] = m[4 * r] * p[0] + m[4 * r + 1] * p[1] + m[4 * r + 2] * p[2] + m[4 * r + 3] * 1.0;
}

struct SceneIn {
  const double* pos;
  const double* nrm;
  const int64_t* off;
  const int32_t* mesh_mat;
  const int32_t* node_mesh;
  const double* matrix;
  std::vector<std::vector<int32_t>> kids;  // [n_nodes] = the root
};

// collectTracables (scene.go:95-114): this node's mesh (raw when its own
// transformation is the identity, else transformed by t), then its children.
void collect(const SceneIn& s, int32_t node, const M4& parent_t, bool is_root, std::vector<double>& pos,
             std::vector<double>& nrm, std::vector<int32_t>& mat) {
  M4 own = m4_identity();
  if (!is_root) std::memcpy(own.m, s.matrix + 16 * size_t(node), sizeof(own.m));
  const M4 t = m4_mul(own, parent_t);
  const int32_t mesh = is_root ? -1 : s.node_mesh[node];
  if (mesh >= 0) {
    bool ident = true;
    const M4 I = m4_identity();
    for (int k = 0; k < 16; ++k) ident = ident && own.m[k] == I.m[k];
    const M4 tinv = m4_inverse(m4_transpose(t));
    for (int64_t i = s.off[mesh]; i < s.off[mesh + 1]; ++i) {
      if (is_sphere(s.pos + 9 * i)) {
        // Sphere.transformed (primitive.go:42-50): the centre as a point, radius kept
        double q[9];
        std::memcpy(q, s.pos + 9 * i, sizeof(q));
        if (!ident) m4_point(t, s.pos + 9 * i, q);
        pos.insert(pos.end(), q, q + 9);
        if (s.nrm) nrm.insert(nrm.end(), 9, 0.0);
        mat.push_back(s.mesh_mat[mesh]);
        continue;
      }
      for (int v = 0; v < 3; ++v) {
        double p[3], n[3];
        const double* sp = s.pos + 9 * i + 3 * v;
        if (ident) {
          std::memcpy(p, sp, sizeof(p));
        } else {
          m4_point(t, sp, p);
        }
        pos.insert(pos.end(), p, p + 3);
        if (s.nrm) {
          const double* sn = s.nrm + 9 * i + 3 * v;
          if (ident) std::memcpy(n, sn, sizeof(n));
          else m4_point(tinv, sn, n);
          nrm.insert(nrm.end(), n, n + 3);
        }
      }
      mat.push_back(s.mesh_mat[mesh]);
    }
  }
  for (int32_t c : s.kids[is_root ? s.kids.size() - 1 : size_t(node)]) collect(s, c, t, false, pos, nrm, mat);
}
}  // namespace

extern "C" {

// Scene.collectTracables from Scene.root (identity, no mesh) over a node tree:
// node k has parent[k] (-1: the root's child), node_mesh[k] (-1: none) and
// matrix[16k ..] (row-major Matrix4).  Writes positions (9 doubles), normals
// (9 doubles, if nrm is given) and material per flattened triangle when the
// outputs are non-NULL; returns the triangle count.
int64_t oracle_scene_flatten(const double* pos, const double* nrm, const int64_t* off, const int32_t* mesh_mat,
                             const int32_t* parent, const int32_t* node_mesh, const double* matrix,
                             int32_t n_nodes, double* out_pos, double* out_nrm, int32_t* out_mat) {
  SceneIn s{pos, nrm, off, mesh_mat, node_mesh, matrix, {}};
  s.kids.resize(size_t(n_nodes) + 1);
  for (int32_t k = 0; k < n_nodes; ++k) s.kids[parent[k] < 0 ? size_t(n_nodes) : size_t(parent[k])].push_back(k);
  std::vector<double> p, n;
  std::vector<int32_t> m;
  collect(s, n_nodes, m4_identity(), true, p, n, m);
  if (out_pos) std::memcpy(out_pos, p.data(), p.size() * sizeof(double));
  if (out_nrm && nrm) std::memcpy(out_nrm, n.data(), n.size() * sizeof(double));
  if (out_mat) std::memcpy(out_mat, m.data(), m.size() * sizeof(int32_t));
  return int64_t(m.size());
}

int64_t oracle_phr_split_work(const oracle_phr* t) { return t->split_work; }
int64_t oracle_phr_initial_cut(const oracle_phr* t) { return t->initial_cut; }
int64_t oracle_phr_max_split(const oracle_phr* t) { return t->max_split; }

// Canonical preorder of the PHR tree with reused LBVH subtrees expanded:
// rec = kind (0 inner, 1 leaf), prim count (leaf) or -1, n_children, 0;
// box FP64; leaf prims (sorted positions) concatenated in preorder.
// Returns the node count; call with null outputs to size the buffers
// (*n_prims receives the prim-list length).
int64_t oracle_phr_canonical(const oracle_phr* t, int32_t* rec, double* box, int32_t* prims,
                             int64_t* n_prims) {
  struct It {
    int32_t id;
    bool aux;
  };
  std::vector<It> st{{t->root, false}};
  int64_t k = 0, np = 0;
  const Tree& aux = *t->aux;
  while (!st.empty()) {
    It it = st.back();
    st.pop_back();
    if (!it.aux && t->nodes[it.id].kind == 2) it = It{t->nodes[it.id].aux, true};
    const Box* b;
    int kind, cnt = -1, nch = 0;
    if (it.aux) {
      const Node& nd = aux.nodes[it.id];
      b = &nd.box;
      if (nd.leaf) {
        kind = 1;
        cnt = nd.last - nd.first + 1;
        for (int32_t i = nd.first; i <= nd.last; ++i) {
          if (prims) prims[np] = i;
          ++np;
        }
      } else {
        kind = 0;
        nch = 2;
        st.push_back({nd.child[1], true});
        st.push_back({nd.child[0], true});
      }
    } else {
      const PNode& p = t->nodes[it.id];
      b = &p.box;
      if (p.kind == 1) {
        kind = 1;
        cnt = int(p.prims.size());
        for (int32_t v : p.prims) {
          if (prims) prims[np] = v;
          ++np;
        }
      } else {
        kind = 0;
        nch = int(p.kids.size());
        for (size_t c = p.kids.size(); c-- > 0;) st.push_back({p.kids[c], false});
      }
    }
    if (rec) {
      rec[4 * k + 0] = kind;
      rec[4 * k + 1] = cnt;
      rec[4 * k + 2] = nch;
      rec[4 * k + 3] = 0;
      for (int a = 0; a < 3; ++a) {
        box[6 * k + a] = b->lo[a];
        box[6 * k + 3 + a] = b->hi[a];
      }
    }
    ++k;
  }
  if (n_prims) *n_prims = np;
  return k;
}

}  // extern "C"

// ---- path tracing (render.go:91-129, 201-246; material.go:13-42) -----------
namespace {

struct Shading {
  const double* nrm;  // 9 per triangle (caller order)
  const int32_t* mat_idx;
  const float* mats;  // kind, r, g, b
  int miss;
  int max_depth;
  uint64_t seed;
  // optional per-triangle RGB albedo (Diffuse) and emission (Light), caller
  // order; they replace the material table's colour
  const float* tri_alb = nullptr;
  const float* tri_emi = nullptr;
  // per material: Reflective.Diffusion / Refractive.Ratio (null: none in use)
  const float* params = nullptr;

  V3 colour(int32_t prim) const {
    const float* m = mats + 4 * mat_idx[prim];
    const float* o = m[0] == 1.0f ? tri_emi : tri_alb;
    if (o) return {o[3 * size_t(prim)], o[3 * size_t(prim) + 1], o[3 * size_t(prim) + 2]};
    return {m[1], m[2], m[3]};
  }
  bool light(int32_t prim) const { return mats[4 * mat_idx[prim]] == 1.0f; }
};

V3 miss_color(int miss, V3 d) {
  switch (miss) {
    case 1:
      return {1, 1, 1};
    case 2: {  // SkyMissShader
      const double len = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
      const double t = 0.5 * (d.y / len + 1);
      return {0.8 * (1.0 - t) + 0.25 * t, 0.8 * (1.0 - t) + 0.35 * t, 0.8 * (1.0 - t) + 0.5 * t};
    }
    case 3:
      return {1, .96, .78};
    default:
      return {0, 0, 0};
  }
}

// BVH.intersected with the hit record the shader needs.
bool trace_ray(const Tree& t, V3 o, V3 d, double tmin, HitRec& h) {
  Ray r;  // camera.go:138-158 ray.reuse
  r.o = o;
  r.d = d;
  r.inv = {1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
  r.sign[0] = r.inv.x < 0;
  r.sign[1] = r.inv.y < 0;
  r.sign[2] = r.inv.z < 0;
  h.t = INFINITY;
  h.prim = -1;
  return node_hit(t, t.root, r, tmin, INFINITY, h);
}

// The hit's interpolated vertex normal, flipped to face the ray
// (primitive.go:145-150, 200-205); not renormalised.
V3 facing_normal(const Shading& s, const HitRec& h) {
  if (h.sphere) return h.n;
  const double* n = s.nrm + 9 * size_t(h.prim);
  const double u = h.u, v = h.v, w0 = 1 - u - v;
  V3 nn{(n[3] * u + n[6] * v) + n[0] * w0, (n[4] * u + n[7] * v) + n[1] * w0,
        (n[5] * u + n[8] * v) + n[2] * w0};
  if (!h.front) nn = {-nn.x, -nn.y, -nn.z};
  return nn;
}

// The uniform unit vector of vertex k (its first two Philox words).
V3 unit_draw(const Philox4& r) {
  const double z = 1.0 - 2.0 * double(u01(r.v[0]));
  const double rad = std::sqrt(std::max(0.0, 1.0 - z * z));
  const double phi = 6.283185307179586 * double(u01(r.v[1]));
  return {rad * std::cos(phi), rad * std::sin(phi), z};
}

// Diffuse.scatter (material.go:29-38): the scattered direction at vertex k.
V3 scatter_dir(const Shading& s, uint64_t path, int k, const HitRec& h) {
  const V3 nn = facing_normal(s, h);
  const V3 w = unit_draw(path_draw(s.seed, path, uint32_t(k + 1)));
  V3 dir{nn.x + w.x, nn.y + w.y, nn.z + w.z};
  if (std::fabs(dir.x) <= 1e-7 && std::fabs(dir.y) <= 1e-7 && std::fabs(dir.z) <= 1e-7) dir = nn;
  return dir;
}

// reflect (material.go:90-92)
V3 reflect_v(V3 v, V3 n) {
  const double k = dot(v, n) * 2;
  return {v.x - n.x * k, v.y - n.y * k, v.z - n.z * k};
}

// Reflective.scatter / Refractive.scatter (material.go:44-106) at vertex k:
// false when the path ends here (a reflection below the surface).
bool specular_dir(const Shading& s, int kind, uint64_t path, int k, V3 d, const HitRec& h, V3& dir) {
  const double prm = s.params ? double(s.params[s.mat_idx[h.prim]]) : 0.0;
  const V3 n = facing_normal(s, h);
  const double len = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);  // math.go Unit
  const V3 unit{d.x / len, d.y / len, d.z / len};
  const Philox4 r = path_draw(s.seed, path, uint32_t(k + 1));
  const V3 refl = reflect_v(unit, n);
  if (kind == 2) {
    const V3 w = unit_draw(r);
    dir = {refl.x + w.x * prm, refl.y + w.y * prm, refl.z + w.z * prm};
    return dot(refl, n) > 0;
  }
  const double ratio = h.front ? 1 / prm : prm;
  const V3 neg{-unit.x, -unit.y, -unit.z};
  const double cos_t = std::min(dot(neg, n), 1.0);
  const double sin_t = std::sqrt(1.0 - cos_t * cos_t);
  double r0 = (1 - ratio) / (1 + ratio);
  r0 = r0 * r0;
  const double reflectance = r0 + (1 - r0) * std::pow(1 - cos_t, 5);  // Schlick
  if (ratio * sin_t > 1.0 || reflectance > double(u01(r.v[2]))) {
    dir = refl;
  } else {
    const V3 perp{(unit.x + n.x * cos_t) * ratio, (unit.y + n.y * cos_t) * ratio, (unit.z + n.z * cos_t) * ratio};
    const double par = -std::sqrt(std::fabs(1.0 - dot(perp, perp)));
    dir = {perp.x + n.x * par, perp.y + n.y * par, perp.z + n.z * par};
  }
  return true;
}

// DefaultClosestHitShader (render.go:209-224): vertex at depth k.
V3 shade(const Tree& t, const Shading& s, uint64_t path, int k, V3 o, V3 d, const HitRec& h) {
  if (k > s.max_depth) return {0, 0, 0};
  const V3 c = s.colour(h.prim);
  if (s.light(h.prim)) return c;  // Light: emission, no scatter
  const int kind = int(s.mats[4 * s.mat_idx[h.prim]]);
  V3 dir;
  if (kind == 0) dir = scatter_dir(s, path, k, h);
  else if (!specular_dir(s, kind, path, k, d, h, dir)) return {0, 0, 0};
  const V3 p{o.x + d.x * h.t, o.y + d.y * h.t, o.z + d.z * h.t};
  HitRec h2{};
  V3 next;
  if (trace_ray(t, p, dir, 0.0001, h2)) next = shade(t, s, path, k + 1, p, dir, h2);
  else next = miss_color(s.miss, dir);
  return {next.x * c.x, next.y * c.y, next.z * c.z};  // Le = 0 for Diffuse
}

// The same path as shade() traced iteratively, keeping its diffuse vertices
// (triangle, albedo) and how it ends: *light = the light triangle it reached
// (-1 for a miss or the depth limit), *E = the radiance there.
void path_tape(const Tree& t, const Shading& s, uint64_t path, V3 o, V3 d,
               std::vector<std::pair<int32_t, V3>>& tape, int32_t* light, V3* E) {
  tape.clear();
  *light = -1;
  double tmin = 0.001;
  for (int k = 0;; ++k) {
    HitRec h{};
    if (!trace_ray(t, o, d, tmin, h)) {
      *E = miss_color(s.miss, d);
      return;
    }
    if (k > s.max_depth) {
      *E = {0, 0, 0};
      return;
    }
    const V3 c = s.colour(h.prim);
    if (s.light(h.prim)) {
      *light = h.prim;
      *E = c;
      return;
    }
    tape.push_back({h.prim, c});
    const V3 dir = scatter_dir(s, path, k, h);
    o = {o.x + d.x * h.t, o.y + d.y * h.t, o.z + d.z * h.t};
    d = dir;
    tmin = 0.0001;
  }
}

}  // namespace

extern "C" {

// geometric vertex normals (primitive.go:116-138) of every triangle
static std::vector<double> geo_normals(const oracle_tree* t) {
  std::vector<double> geo(9 * size_t(t->n));
  for (int64_t i = 0; i < t->n; ++i) {
    const double* p = &t->pos[9 * size_t(i)];
    if (is_sphere(p)) continue;  // a sphere's normal comes from its centre (zeros here)
    for (int c = 0; c < 3; ++c) {
      const V3 a{p[3 * c], p[3 * c + 1], p[3 * c + 2]};
      const int ri = (c + 1) % 3, li = (c + 2) % 3;
      const V3 rgt{p[3 * ri], p[3 * ri + 1], p[3 * ri + 2]};
      const V3 lft{p[3 * li], p[3 * li + 1], p[3 * li + 2]};
      const V3 cr = cross(sub(rgt, a), sub(lft, a));
      const double len = std::sqrt(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
      geo[9 * i + 3 * c + 0] = cr.x / len;
      geo[9 * i + 3 * c + 1] = cr.y / len;
      geo[9 * i + 3 * c + 2] = cr.z / len;
    }
  }
  return geo;
}

// camera ray of (pixel, path) (camera.go:95-97, render.go:201-205)
static V3 camera_dir(const double* cam, int32_t W, int32_t H, int64_t pix, uint64_t seed, uint64_t path) {
  const int32_t x = int32_t(pix % W), y = int32_t(pix / W);
  const Philox4 r = path_draw(seed, path, 0u);
  const double sx = (double(x) + double(u01(r.v[0]))) / double(W - 1);
  const double sy = (double(y) + double(u01(r.v[1]))) / double(H - 1);
  return {cam[3] + cam[6] * sx + cam[9] * sy - cam[0], cam[4] + cam[7] * sx + cam[10] * sy - cam[1],
          cam[5] + cam[8] * sx + cam[11] * sy - cam[2]};
}

// Per-pixel sum, sum of squares and last sample of spp paths, FP64.  cam =
// origin, lower_left, horizontal, vertical (12 doubles).  nrm = NULL uses
// geometric normals (primitive.go:116-138).  tri_alb / tri_emi (optional, 3
// floats per triangle) replace the material colour of Diffuse / Light
// triangles.  path_out (optional) receives every path's value, indexed
// [sample][pixel][rgb].
static void render_impl(const oracle_tree* t, const double* nrm, const int32_t* mat_idx, const float* mats,
                        const float* params, const float* tri_alb, const float* tri_emi, const double* cam,
                        int32_t W, int32_t H, int32_t spp, int32_t max_depth, int32_t miss, uint64_t seed,
                        int64_t sample_offset, int threads, double* sum, double* sumsq, double* last,
                        double* path_out) {
  std::vector<double> geo;
  if (!nrm) {
    geo = geo_normals(t);
    nrm = geo.data();
  }
  Shading s{nrm, mat_idx, mats, miss, max_depth, seed};
  s.tri_alb = tri_alb;
  s.tri_emi = tri_emi;
  s.params = params;
  const V3 org{cam[0], cam[1], cam[2]};
  const int64_t wh = int64_t(W) * H;
  parallel_for(wh, threads, [&](int64_t b, int64_t e) {
    for (int64_t pix = b; pix < e; ++pix) {
      double S[3] = {0, 0, 0}, Q[3] = {0, 0, 0};
      V3 L{0, 0, 0};
      for (int32_t smp = 0; smp < spp; ++smp) {
        const uint64_t path = uint64_t((sample_offset + smp) * wh + pix);
        const V3 d = camera_dir(cam, W, H, pix, seed, path);
        HitRec h{};
        if (trace_ray(*t, org, d, 0.001, h)) L = shade(*t, s, path, 0, org, d, h);
        else L = miss_color(miss, d);
        S[0] += L.x; S[1] += L.y; S[2] += L.z;
        Q[0] += L.x * L.x; Q[1] += L.y * L.y; Q[2] += L.z * L.z;
        if (path_out) {
          double* po = path_out + 3 * (size_t(smp) * wh + pix);
          po[0] = L.x; po[1] = L.y; po[2] = L.z;
        }
      }
      for (int c = 0; c < 3; ++c) {
        sum[3 * pix + c] = S[c];
        sumsq[3 * pix + c] = Q[c];
      }
      if (last) {
        last[3 * pix + 0] = L.x;
        last[3 * pix + 1] = L.y;
        last[3 * pix + 2] = L.z;
      }
    }
  });
}

void oracle_render(const oracle_tree* t, const double* nrm, const int32_t* mat_idx,
                   const float* mats, const float* tri_alb, const float* tri_emi, const double* cam,
                   int32_t W, int32_t H, int32_t spp, int32_t max_depth, int32_t miss, uint64_t seed,
                   int64_t sample_offset, int threads, double* sum, double* sumsq, double* last,
                   double* path_out) {
  render_impl(t, nrm, mat_idx, mats, nullptr, tri_alb, tri_emi, cam, W, H, spp, max_depth, miss, seed,
              sample_offset, threads, sum, sumsq, last, path_out);
}

// The same with all four material kinds (material.go:13-106): 8 floats per
// material, kind, r, g, b, then Reflective.Diffusion / Refractive.Ratio.
void oracle_render_ex(const oracle_tree* t, const double* nrm, const int32_t* mat_idx, const float* mats8,
                      int32_t nmat, const double* cam, int32_t W, int32_t H, int32_t spp, int32_t max_depth,
                      int32_t miss, uint64_t seed, int threads, double* sum, double* sumsq, double* last) {
  std::vector<float> mats(4 * size_t(nmat)), params(nmat);
  for (int32_t m = 0; m < nmat; ++m) {
    for (int c = 0; c < 4; ++c) mats[4 * size_t(m) + c] = mats8[8 * size_t(m) + c];
    params[m] = mats8[8 * size_t(m) + 4];
  }
  render_impl(t, nrm, mat_idx, mats.data(), params.data(), nullptr, nullptr, cam, W, H, spp, max_depth, miss, seed, 0,
              threads, sum, sumsq, last, nullptr);
}

// Gradient of a loss on the mean image (sum / spp) of the render above with
// respect to the per-triangle albedo and emission, given g_img = dLoss/d(mean
// image) (W*H*3).  Each path's value is L = a_0 (x) a_1 (x) ... (x) a_{m-1} (x) E
// (diffuse albedos along the path, E the light's emission or the miss colour),
// and the albedo does not steer the path, so
//   dL/da_j = (prod_{i<j} a_i) (x) (prod_{i>j} a_i) (x) E   (summed over repeats)
//   dL/dEmission(light) = prod_i a_i
// evaluated with explicit prefix and suffix products (no division).  Also
// writes the mean image.  Outputs are FP64; g_alb and g_emi hold 3 per
// triangle, caller order, and are overwritten.
void oracle_render_grad(const oracle_tree* t, const double* nrm, const int32_t* mat_idx,
                        const float* mats, const float* tri_alb, const float* tri_emi,
                        const double* cam, int32_t W, int32_t H, int32_t spp, int32_t max_depth,
                        int32_t miss, uint64_t seed, int64_t sample_offset, int threads,
                        const double* g_img, double* mean, double* g_alb, double* g_emi) {
  std::vector<double> geo;
  if (!nrm) {
    geo = geo_normals(t);
    nrm = geo.data();
  }
  Shading s{nrm, mat_idx, mats, miss, max_depth, seed};
  s.tri_alb = tri_alb;
  s.tri_emi = tri_emi;
  const V3 org{cam[0], cam[1], cam[2]};
  const int64_t wh = int64_t(W) * H;
  const int64_t n3 = 3 * t->n;
  // one accumulator per parallel_for chunk, summed in chunk order afterwards
  const bool serial = threads <= 1 || wh < 1024;
  const int nchunk = serial ? 1 : threads;
  const int64_t chunk = serial ? wh : (wh + threads - 1) / threads;
  std::vector<std::vector<double>> acc(nchunk);
  parallel_for(wh, threads, [&](int64_t b, int64_t e) {
    std::vector<double>& A = acc[size_t(b / chunk)];
    A.assign(2 * size_t(n3), 0.0);
    double* ga = A.data();
    double* ge = A.data() + n3;
    std::vector<std::pair<int32_t, V3>> tape;
    std::vector<V3> suffix;
    for (int64_t pix = b; pix < e; ++pix) {
      const double w[3] = {g_img[3 * pix] / spp, g_img[3 * pix + 1] / spp, g_img[3 * pix + 2] / spp};
      double S[3] = {0, 0, 0};
      for (int32_t smp = 0; smp < spp; ++smp) {
        const uint64_t path = uint64_t((sample_offset + smp) * wh + pix);
        const V3 d = camera_dir(cam, W, H, pix, seed, path);
        int32_t light;
        V3 E;
        path_tape(*t, s, path, org, d, tape, &light, &E);
        const size_t m = tape.size();
        suffix.resize(m + 1);
        suffix[m] = E;  // radiance arriving from beyond vertex j: prod_{i>j} a_i (x) E
        for (size_t j = m; j-- > 0;) {
          const V3 a = tape[j].second;
          suffix[j] = {suffix[j + 1].x * a.x, suffix[j + 1].y * a.y, suffix[j + 1].z * a.z};
        }
        S[0] += suffix[0].x;
        S[1] += suffix[0].y;
        S[2] += suffix[0].z;
        V3 pre{1, 1, 1};
        for (size_t j = 0; j < m; ++j) {
          const int32_t tri = tape[j].first;
          ga[3 * tri + 0] += w[0] * (pre.x * suffix[j + 1].x);
          ga[3 * tri + 1] += w[1] * (pre.y * suffix[j + 1].y);
          ga[3 * tri + 2] += w[2] * (pre.z * suffix[j + 1].z);
          const V3 a = tape[j].second;
          pre = {pre.x * a.x, pre.y * a.y, pre.z * a.z};
        }
        if (light >= 0) {
          ge[3 * light + 0] += w[0] * pre.x;
          ge[3 * light + 1] += w[1] * pre.y;
          ge[3 * light + 2] += w[2] * pre.z;
        }
      }
      for (int c = 0; c < 3; ++c) mean[3 * pix + c] = S[c] / spp;
    }
  });
  for (int64_t i = 0; i < n3; ++i) {
    double a = 0, em = 0;
    for (const auto& A : acc) {
      if (A.empty()) continue;
      a += A[size_t(i)];
      em += A[size_t(n3 + i)];
    }
    g_alb[i] = a;
    g_emi[i] = em;
  }
}

}  // extern "C"
