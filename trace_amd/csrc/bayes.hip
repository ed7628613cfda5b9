// Bayesian search of the PHR parameters: bayesianOptimizer.OptimizedPHRparams
// (optimizer.go:96-155).
//
// The reference evaluates two fixed builds for the normalisers -- Cost() of
// the tree at (alpha_min, delta_min) and the build cost at (alpha_max,
// delta_max) -- and then hands the objective
//   evalPHR = buildCost / maxBuildCost + omega^2 * Cost / maxCost
// to the third-party module changkun.de/x/bo v0.0.1 (bo.NewOptimizer with
// minimisation and expected improvement, RunSerial), which is not part of the
// reference tree.  The objective here is the reference's, evaluated with the
// device PHR build and Cost(); the search is this library's own Gaussian-process
// optimiser of the same published kind:
//   - the first rounds sample the box uniformly (a seeded splitmix64 stream);
//   - every later round fits a GP (Matern 5/2 kernel in unit-cube coordinates,
//     standardised observations, the length scale of largest marginal
//     likelihood among a few, Cholesky with growing jitter) and evaluates the
//     point of largest expected improvement, found over random and lattice
//     candidates and refined by a compass search;
//   - the result is the evaluated point of smallest objective (first minimum).
// The search is host code: one objective evaluation (a device PHR build and
// Cost()) takes milliseconds, the GP algebra over a few dozen points
// microseconds.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

namespace {

struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double unit() { return double(next() >> 11) * 0x1.0p-53; }  // [0, 1)
};

double matern52(const double* a, const double* b, int dims, double len) {
  double d2 = 0.0;
  for (int k = 0; k < dims; ++k) {
    const double d = (a[k] - b[k]) / len;
    d2 += d * d;
  }
  const double r = std::sqrt(5.0 * d2);
  return (1.0 + r + r * r / 3.0) * std::exp(-r);
}

// GP regression on n points u (n x dims, unit cube) with observations y.
struct Gp {
  int dims = 0, n = 0;
  double len = 0.25, ymean = 0.0, yscale = 1.0, best = 0.0;
  std::vector<double> u, L, alpha;

  // Cholesky of K + jitter I in place; false when not positive definite.
  static bool cholesky(std::vector<double>& A, int n) {
    for (int j = 0; j < n; ++j) {
      double d = A[j * n + j];
      for (int k = 0; k < j; ++k) d -= A[j * n + k] * A[j * n + k];
      if (!(d > 0.0)) return false;
      d = std::sqrt(d);
      A[j * n + j] = d;
      for (int i = j + 1; i < n; ++i) {
        double s = A[i * n + j];
        for (int k = 0; k < j; ++k) s -= A[i * n + k] * A[j * n + k];
        A[i * n + j] = s / d;
      }
      for (int i = 0; i < j; ++i) A[i * n + j] = 0.0;
    }
    return true;
  }
  void solve_lower(const double* b, double* x) const {  // L x = b
    for (int i = 0; i < n; ++i) {
      double s = b[i];
      for (int k = 0; k < i; ++k) s -= L[i * n + k] * x[k];
      x[i] = s / L[i * n + i];
    }
  }
  void solve_upper(const double* b, double* x) const {  // L^T x = b
    for (int i = n - 1; i >= 0; --i) {
      double s = b[i];
      for (int k = i + 1; k < n; ++k) s -= L[k * n + i] * x[k];
      x[i] = s / L[i * n + i];
    }
  }
  // factorises at length scale l; returns the log marginal likelihood (-inf on failure)
  double factor(const std::vector<double>& ys, double l) {
    len = l;
    for (double jitter = 1e-8; jitter <= 1e-2; jitter *= 100.0) {
      L.assign(size_t(n) * n, 0.0);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
          const double k = matern52(&u[i * dims], &u[j * dims], dims, len) + (i == j ? jitter : 0.0);
          L[i * n + j] = k;
          L[j * n + i] = k;
        }
      if (!cholesky(L, n)) continue;
      std::vector<double> t(n);
      alpha.assign(n, 0.0);
      solve_lower(ys.data(), t.data());
      solve_upper(t.data(), alpha.data());
      double fit = 0.0, logdet = 0.0;
      for (int i = 0; i < n; ++i) {
        fit += ys[i] * alpha[i];
        logdet += std::log(L[i * n + i]);
      }
      return -0.5 * fit - logdet;
    }
    return -INFINITY;
  }
  bool fit(const std::vector<double>& pts, const std::vector<double>& y, int d) {
    dims = d;
    n = int(y.size());
    u = pts;
    ymean = 0.0;
    for (double v : y) ymean += v;
    ymean /= n;
    double var = 0.0;
    for (double v : y) var += (v - ymean) * (v - ymean);
    yscale = n > 1 ? std::sqrt(var / n) : 0.0;
    if (!(yscale > 1e-300)) yscale = 1.0;
    std::vector<double> ys(n);
    best = INFINITY;
    for (int i = 0; i < n; ++i) {
      ys[i] = (y[i] - ymean) / yscale;
      best = std::min(best, ys[i]);
    }
    // length scale of largest marginal likelihood (first maximum)
    static const double kLens[] = {0.05, 0.1, 0.2, 0.35, 0.6, 1.0};
    double best_ll = -INFINITY, best_len = -1.0;
    for (double l : kLens) {
      const double ll = factor(ys, l);
      if (ll > best_ll) {
        best_ll = ll;
        best_len = l;
      }
    }
    if (best_len < 0.0) return false;
    return factor(ys, best_len) > -INFINITY;
  }
  // expected improvement below the best standardised observation
  double ei(const double* x) const {
    std::vector<double> k(n), v(n);
    double mean = 0.0;
    for (int i = 0; i < n; ++i) {
      k[i] = matern52(x, &u[i * dims], dims, len);
      mean += k[i] * alpha[i];
    }
    solve_lower(k.data(), v.data());
    double var = 1.0;
    for (int i = 0; i < n; ++i) var -= v[i] * v[i];
    const double sd = var > 0.0 ? std::sqrt(var) : 0.0;
    const double gain = best - mean;
    if (sd < 1e-12) return std::max(gain, 0.0);
    const double z = gain / sd;
    const double cdf = 0.5 * std::erfc(-z / std::sqrt(2.0));
    const double pdf = std::exp(-0.5 * z * z) / std::sqrt(2.0 * M_PI);
    return gain * cdf + sd * pdf;
  }
};

// The point of largest expected improvement: random and lattice candidates,
// then a compass search from the best of them.
void next_point(const Gp& gp, int dims, SplitMix& rng, double* out) {
  std::vector<double> x(dims), best_x(dims);
  double best = -INFINITY;
  auto consider = [&](const double* c) {
    const double e = gp.ei(c);
    if (e > best) {
      best = e;
      std::copy(c, c + dims, best_x.begin());
    }
  };
  for (int c = 0; c < 1024; ++c) {
    for (int k = 0; k < dims; ++k) x[k] = rng.unit();
    consider(x.data());
  }
  if (dims <= 3) {
    const int side = 17;
    int total = 1;
    for (int k = 0; k < dims; ++k) total *= side;
    for (int c = 0; c < total; ++c) {
      int r = c;
      for (int k = 0; k < dims; ++k) {
        x[k] = double(r % side) / double(side - 1);
        r /= side;
      }
      consider(x.data());
    }
  }
  double step = 1.0 / 16.0;
  for (int it = 0; it < 64 && step > 1.0 / 4096.0; ++it) {
    bool moved = false;
    for (int k = 0; k < dims && !moved; ++k)
      for (int s = -1; s <= 1 && !moved; s += 2) {
        x = best_x;
        x[k] = std::min(1.0, std::max(0.0, x[k] + s * step));
        const double e = gp.ei(x.data());
        if (e > best) {
          best = e;
          best_x = x;
          moved = true;
        }
      }
    if (!moved) step *= 0.5;
  }
  std::copy(best_x.begin(), best_x.end(), out);
}

struct PhrObjective {
  trace_ctx* ctx;
  const trace_bvh* aux;
  int branching;
  int32_t max_cut;
  double max_sah;
  int64_t max_build;
  double omega;
  std::vector<double> rows;  // alpha, delta, Cost, build cost, objective
  trace_bvh* tree = nullptr;  // every evaluation is rebuilt into the first one's storage
  ~PhrObjective() {
    if (tree) trace_bvh_destroy(tree);
  }
};

int phr_eval(PhrObjective* o, double a, double d, double* sah, int64_t* build) {
  int rc = o->tree ? trace_phr_rebuild(o->ctx, o->aux, a, d, o->branching, o->max_cut, o->tree)
                   : trace_phr_build(o->ctx, o->aux, a, d, o->branching, o->max_cut, &o->tree);
  if (rc) return rc;
  *build = o->tree->phr_initial_cut + o->tree->phr_split_work;
  return trace_bvh_cost(o->ctx, o->tree, sah);
}

// evalPHR (optimizer.go:92-94) of one (alpha, delta), with the reference's
// substitute for a zero parameter (optimizer.go:142-144)
int phr_objective(void* user, const double* x, double* y) {
  auto* o = static_cast<PhrObjective*>(user);
  const double a = x[0], d = x[1];
  double sah = o->max_sah;
  int64_t build = o->max_build;
  if (a != 0.0 && d != 0.0) {
    const int rc = phr_eval(o, a, d, &sah, &build);
    if (rc) return rc;
  }
  *y = double(build) / double(o->max_build) + (o->omega * o->omega) * (sah / o->max_sah);
  o->rows.insert(o->rows.end(), {a, d, sah, double(build), *y});
  return TRACE_OK;
}

}  // namespace

extern "C" {

int trace_bayes_minimize(const double* lo, const double* hi, int32_t dims, int32_t rounds, int32_t random_rounds,
                         uint64_t seed, trace_objective_fn objective, void* user, double* h_x, double* h_y,
                         double* x_best) {
  if (!lo || !hi || !objective || !x_best || dims < 1 || dims > 8 || rounds < 1 || random_rounds < 0)
    return TRACE_E_INVALID;
  for (int k = 0; k < dims; ++k)
    if (!(lo[k] <= hi[k]) || !std::isfinite(lo[k]) || !std::isfinite(hi[k])) return TRACE_E_INVALID;
  SplitMix rng{seed};
  std::vector<double> pts, ys;  // unit-cube points and objectives, in evaluation order
  std::vector<double> unit(dims), x(dims);
  int best_i = -1;
  for (int r = 0; r < rounds; ++r) {
    Gp gp;
    const bool model = r >= std::max(1, int(random_rounds)) && gp.fit(pts, ys, dims);
    if (model) {
      next_point(gp, dims, rng, unit.data());
    } else {
      for (int k = 0; k < dims; ++k) unit[k] = rng.unit();
    }
    for (int k = 0; k < dims; ++k) x[k] = lo[k] + unit[k] * (hi[k] - lo[k]);
    double y = 0.0;
    const int rc = objective(user, x.data(), &y);
    if (rc) return rc;
    if (std::isnan(y)) return TRACE_E_INVALID;
    pts.insert(pts.end(), unit.begin(), unit.end());
    ys.push_back(y);
    if (h_x) std::copy(x.begin(), x.end(), h_x + size_t(r) * dims);
    if (h_y) h_y[r] = y;
    if (best_i < 0 || y < ys[best_i]) {
      best_i = r;
      std::copy(x.begin(), x.end(), x_best);
    }
  }
  return TRACE_OK;
}

int trace_phr_optimize_bayes(trace_ctx* ctx, const trace_bvh* aux, const double* alpha_range,
                             const double* delta_range, int branching, int32_t max_cut, int32_t frame_width,
                             int32_t frame_height, int32_t rounds, int32_t random_rounds, uint64_t seed,
                             double* h_trace, double* alpha_out, double* delta_out) {
  if (!ctx) return TRACE_E_INVALID;
  if (!aux || !alpha_range || !delta_range || frame_width < 1 || frame_height < 1 || rounds < 1 ||
      random_rounds < 0 || !alpha_out || !delta_out || !(alpha_range[0] <= alpha_range[1]) ||
      !(delta_range[0] <= delta_range[1]))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_optimize_bayes: bad arguments");
  PhrObjective o{ctx, aux, branching, max_cut, 0.0, 0, 0.0, {}, nullptr};
  // normalisers (optimizer.go:136-140): Cost() at the range minimum, the build
  // cost at the range maximum
  int64_t build_lo = 0;
  double sah_hi = 0.0;
  int rc = phr_eval(&o, alpha_range[0], delta_range[0], &o.max_sah, &build_lo);
  if (rc) return rc;
  rc = phr_eval(&o, alpha_range[1], delta_range[1], &sah_hi, &o.max_build);
  if (rc) return rc;
  // evalPHR divides by both normalisers; the reference's objective is NaN
  // everywhere when either is zero, so the search is refused up front
  if (o.max_build <= 0 || !(o.max_sah > 0.0))
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_phr_optimize_bayes: zero build or SAH cost at the range ends");
  o.rows.insert(o.rows.end(), {alpha_range[0], delta_range[0], o.max_sah, double(build_lo), NAN});
  o.rows.insert(o.rows.end(), {alpha_range[1], delta_range[1], sah_hi, double(o.max_build), NAN});
  const double r = double(aux->n_prims) / 2000000.0;  // omega (optimizer.go:88-90)
  o.omega = (r * r) * (double(int64_t(frame_width) * frame_height) / 200000.0);
  const double lo[2] = {alpha_range[0], delta_range[0]};
  const double hi[2] = {alpha_range[1], delta_range[1]};
  double best[2] = {0.0, 0.0};
  rc = trace_bayes_minimize(lo, hi, 2, rounds, random_rounds, seed, phr_objective, &o, nullptr, nullptr, best);
  if (rc) {
    if (rc == TRACE_E_INVALID) TRACE_FAIL(ctx, rc, "trace_phr_optimize_bayes: the search failed");
    return rc;  // the failing build set the message
  }
  *alpha_out = best[0];
  *delta_out = best[1];
  if (h_trace) std::copy(o.rows.begin(), o.rows.end(), h_trace);
  return TRACE_OK;
}

}  // extern "C"
