/*
 * CPU oracle for the fastEDM prediction core — TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * A restatement of the reference CPU path of EDM-Developers/fastEDM-r.  Every function cites the
 * reference file:line it follows.  Build: -O2 -ffp-contract=off (no FMA contraction, as the
 * reference's default R build on x86-64).  Nothing here uses Eigen; the Jacobi SVD is a restatement
 * of Eigen 3.4's published JacobiSVD algorithm (the reference calls it at src/edm.cpp:780-782).
 */
#include "oracle.h"

extern "C" double orc_weight_exp(double x);

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <random>
#include <string>
#include <thread>
#include <vector>

namespace orc {

const double MISSING_D = 1.0e+100;
const int MISSING_I = INT_MIN;

enum
{
  SUCCESS = 0,
  BREAK_HIT = 1,
  INVALID_ALGORITHM = 400,
  INSUFFICIENT_UNIQUE = 503,
  UNKNOWN_ERROR = 8000
};

/* ---------------------------------------------------------------- statistics (src/stats.cpp) */

// src/stats.cpp:10-28 — the value of the median order statistic(s).
static double median(std::vector<double> u)
{
  size_t n = u.size();
  if (n % 2 == 0) {
    std::nth_element(u.begin(), u.begin() + n / 2 - 1, u.end());
    double e1 = u[n / 2 - 1];
    std::nth_element(u.begin(), u.begin() + n / 2, u.end());
    double e2 = u[n / 2];
    return (e1 + e2) / 2;
  }
  std::nth_element(u.begin(), u.begin() + n / 2, u.end());
  return u[n / 2];
}

// src/stats.cpp:30-47 — 1-based rank by (value, position).
static std::vector<int> rank_of(const std::vector<double>& v)
{
  std::vector<std::pair<double, int>> s(v.size());
  for (size_t i = 0; i < v.size(); i++)
    s[i] = { v[i], (int)i };
  std::sort(s.begin(), s.end());
  std::vector<int> r(v.size());
  for (size_t i = 0; i < s.size(); i++)
    r[s[i].second] = (int)i + 1;
  return r;
}

static double mean_seq(const double* a, size_t n)
{
  double s = 0.0;
  for (size_t i = 0; i < n; i++)
    s += a[i];
  return s / (double)n;
}

// src/stats.cpp:64-82 — two-pass centred Pearson correlation, MISSING when either SD < 1e-9.
static double correlation(const std::vector<double>& y1, const std::vector<double>& y2)
{
  size_t n = y1.size();
  double m1 = mean_seq(y1.data(), n), m2 = mean_seq(y2.data(), n);
  double s11 = 0, s22 = 0, s12 = 0;
  for (size_t i = 0; i < n; i++) {
    double a = y1[i] - m1, b = y2[i] - m2;
    s11 += a * a;
    s22 += b * b;
  }
  double sd1 = std::sqrt(s11), sd2 = std::sqrt(s22);
  if (sd1 < 1e-9 || sd2 < 1e-9)
    return MISSING_D;
  for (size_t i = 0; i < n; i++)
    s12 += (y1[i] - m1) * (y2[i] - m2);
  return s12 / (sd1 * sd2);
}

// src/stats.cpp:84-94 — mean absolute error, 0 below 1e-8.
static double mean_absolute_error(const std::vector<double>& y1, const std::vector<double>& y2)
{
  double s = 0;
  for (size_t i = 0; i < y1.size(); i++)
    s += std::fabs(y1[i] - y2[i]);
  double mae = s / (double)y1.size();
  return mae < 1e-8 ? 0 : mae;
}

// src/stats.cpp:96-101 — sample standard deviation.
static double standard_deviation(const std::vector<double>& v)
{
  double m = mean_seq(v.data(), v.size());
  double s = 0;
  for (double a : v)
    s += (a - m) * (a - m);
  return std::sqrt(s / (double)(v.size() - 1));
}

static std::vector<double> drop_missing(const std::vector<double>& v)
{
  std::vector<double> o;
  for (double a : v)
    if (a != MISSING_D)
      o.push_back(a);
  return o;
}

// src/stats.cpp:103-108
static double default_missing_distance(const std::vector<double>& x)
{
  return 2 / std::sqrt(M_PI) * standard_deviation(drop_missing(x));
}

/* ---------------------------------------------------------------- MT19937-64 (vendor/mersennetwister.h) */

// src/vendor/mersennetwister.h:121-146.  After init(seed) the first getReal2() returns the tempered
// seed itself (left_ starts at 1, post-decrement at :139); afterwards the stream is exactly the
// standard MT19937-64 stream, which std::mt19937_64 implements.
struct Mt64
{
  std::mt19937_64 eng;
  bool first = true;
  uint64_t seed = 0;
  void init(uint64_t s)
  {
    seed = s;
    eng.seed(s);
    first = true;
  }
  static uint64_t temper(uint64_t y)
  {
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= (y >> 43);
    return y;
  }
  double real2()
  {
    uint64_t v;
    if (first) {
      first = false;
      v = temper(seed);
    } else {
      v = eng();
    }
    return (double)(v >> 11) * (1.0 / 9007199254740992.0);
  }
};

/* ---------------------------------------------------------------- manifold (src/manifold.*) */

struct Generator
{
  std::vector<double> t, x, xmap, co_x;
  std::vector<std::vector<double>> extras;
  std::vector<int> panelIDs, obs;
  int tau = 1, p = 0, num_extras = 0, num_extras_lagged = 0;
  bool dt = false, reldt = false, allow_missing = false, panel_mode = false, xmap_mode = false;
  double dtWeight = 0;

  // src/manifold.cpp:7-18
  static double gcd(double a, double b)
  {
    for (;;) {
      if (a < b)
        std::swap(a, b);
      if (std::fabs(b) < 0.001)
        return a;
      double r = a - std::floor(a / b) * b;
      a = b;
      b = r;
    }
  }

  // src/manifold.cpp:20-48
  double time_increment() const
  {
    double unit = -1;
    for (size_t i = 1; i < t.size(); i++) {
      double d = t[i] - t[i - 1];
      if (d <= 0 || t[i] == MISSING_D || t[i - 1] == MISSING_D)
        continue;
      if (unit < 0) {
        unit = d;
        continue;
      }
      unit = gcd(d, unit);
    }
    return unit;
  }

  // src/manifold.cpp:50-80
  void setup_observation_numbers()
  {
    obs.clear();
    if (!(dt || reldt)) {
      double unit = time_increment();
      double minT = *std::min_element(t.begin(), t.end());
      for (size_t i = 0; i < t.size(); i++)
        obs.push_back(t[i] != MISSING_D ? (int)std::round((t[i] - minT) / unit) : MISSING_I);
    } else {
      int c = 0;
      for (size_t i = 0; i < t.size(); i++) {
        if (t[i] != MISSING_D && (allow_missing || x[i] != MISSING_D))
          obs.push_back(c++);
        else
          obs.push_back(MISSING_I);
      }
    }
  }

  static int wrap_add(int a, long long b) { return (int)(uint32_t)((uint32_t)a + (uint32_t)(int64_t)b); }

  // src/manifold.cpp:82-123 — bounded walk from k in steps of `dir`.
  bool find_obs(int target, int& k, int dir, int panel) const
  {
    while (k >= 0 && k < (int)obs.size()) {
      if (panel_mode) {
        if (panelIDs[k] == MISSING_I) {
          k += dir;
          continue;
        }
        if (panel != panelIDs[k])
          return false;
      }
      if (obs[k] == MISSING_I) {
        k += dir;
        continue;
      }
      if (obs[k] == target)
        return true;
      if (dir > 0 && obs[k] > target)
        return false;
      if (dir < 0 && obs[k] < target)
        return false;
      k += dir;
    }
    return false;
  }

  int E_dt(int E) const { return (dt || reldt) * E; }
  int E_extras(int E) const { return num_extras + num_extras_lagged * (E - 1); }
  int E_actual(int E) const { return E + E_dt(E) + E_extras(E); }

  // src/manifold.cpp:281-312
  double get_target(int i, bool copred, bool predSet, int& ti) const
  {
    int panel = panel_mode ? panelIDs[i] : -1;
    bool use_co_x = copred && predSet;
    if (p != 0) {
      int target = wrap_add(obs[ti], p);
      if (!find_obs(target, ti, p > 0 ? 1 : -1, panel))
        ti = -1;
    }
    if (ti < 0)
      return MISSING_D;
    if (use_co_x)
      return co_x[ti];
    if (xmap_mode)
      return xmap[ti];
    return x[ti];
  }

  // src/manifold.cpp:240-279
  double get_dt(int i) const
  {
    double tNow = t[i];
    if (tNow == MISSING_D)
      return MISSING_D;
    int target = wrap_add(obs[i], -(long long)tau);
    int li = i - 1;
    int panel = panel_mode ? panelIDs[i] : -1;
    if (!find_obs(target, li, -1, panel))
      return MISSING_D;
    double tPrev = t[li];
    if (!reldt)
      return tNow - tPrev;
    int ti = i;
    get_target(i, false, true, ti);
    if (ti < 0)
      return MISSING_D;
    return t[ti] - tPrev;
  }

  // src/manifold.cpp:156-238 — point layout [x lags | dt lags | lagged extras | unlagged extras].
  void fill_in_point(int i, int E, bool copred, bool predSet, double w_dt, double* point) const
  {
    int panel = panel_mode ? panelIDs[i] : -1;
    bool use_co_x = copred && predSet;
    double tPred = MISSING_D;
    if (dt || reldt) {
      int ti = i;
      get_target(i, copred, predSet, ti);
      tPred = ti >= 0 ? t[ti] : MISSING_D;
    }
    int lagged = i, start = obs[i], k = i - 1, prevLagged = lagged;
    int Edt = E_dt(E);
    for (int j = 0; j < E; j++) {
      bool found = (j == 0);
      if (!found) {
        int target = wrap_add(start, -(long long)j * tau);
        if (find_obs(target, k, -1, panel)) {
          found = true;
          lagged = k;
        }
      }
      if (!found)
        point[j] = MISSING_D;
      else
        point[j] = use_co_x ? co_x[lagged] : x[lagged];

      if (dt || reldt) {
        if (!found) {
          point[j + E] = MISSING_D;
        } else {
          double tNow = t[lagged];
          if (j == 0 || reldt) {
            point[E + j] = (tNow != MISSING_D && tPred != MISSING_D) ? w_dt * (tPred - tNow) : MISSING_D;
          } else {
            double tNext = t[prevLagged];
            point[E + j] = (tNext != MISSING_D && tNow != MISSING_D) ? w_dt * (tNext - tNow) : MISSING_D;
          }
        }
      }
      for (int e = 0; e < num_extras; e++) {
        double v = found ? extras[e][lagged] : MISSING_D;
        if (e < num_extras_lagged)
          point[E + Edt + e * E + j] = v;
        else if (j == 0)
          point[E + Edt + num_extras_lagged * E + (e - num_extras_lagged)] = v;
      }
      prevLagged = lagged;
    }
  }

  // src/manifold.h:53-67
  double default_dt_weight()
  {
    double xSD = standard_deviation(drop_missing(x));
    std::vector<double> dts;
    for (size_t i = 0; i < t.size(); i++)
      dts.push_back(get_dt((int)i));
    double dtSD = standard_deviation(drop_missing(dts));
    if (dtSD == 0.0)
      return -1;
    return xSD / dtSD;
  }

  // src/manifold.h:85-127
  void init(bool dt_, bool reldt_, double w)
  {
    panel_mode = panelIDs.size() > 0;
    xmap_mode = xmap.size() > 0;
    num_extras = (int)extras.size();
    dt = dt_;
    reldt = reldt_;
    dtWeight = w;
    if ((dt_ || reldt_) && w == 0.0) {
      dt = true;
      reldt = false;
      setup_observation_numbers();
      double d = default_dt_weight();
      if (d > 0) {
        dt = dt_;
        reldt = reldt_;
        dtWeight = d;
      } else {
        dt = false;
        reldt = false;
        dtWeight = 0;
      }
    }
    setup_observation_numbers();
  }

  // src/manifold.cpp:314-360
  std::vector<bool> generate_usable(int maxE, bool copred) const
  {
    std::vector<bool> usable(t.size());
    int sz = E_actual(maxE);
    std::vector<double> pt(sz);
    for (size_t i = 0; i < t.size(); i++) {
      if (t[i] == MISSING_D || (panel_mode && panelIDs[i] == MISSING_I)) {
        usable[i] = false;
        continue;
      }
      fill_in_point((int)i, maxE, copred, copred, 1.0, pt.data());
      bool ok;
      if (allow_missing) {
        ok = false;
        for (double v : pt)
          if (v != MISSING_D) {
            ok = true;
            break;
          }
      } else {
        ok = true;
        for (double v : pt)
          if (v == MISSING_D) {
            ok = false;
            break;
          }
      }
      usable[i] = ok;
    }
    return usable;
  }
};

// src/manifold.h:159-286 (eager storage; the lazy mode produces the same values).
struct Manifold
{
  std::vector<double> flat, targets, targetTimes, pointTimes;
  std::vector<int> panelIDs, rows;
  int n = 0, E_x = 0, E_dt = 0, E_extras = 0, E_lagged_extras = 0, E_actual = 0;

  Manifold(const Generator& g, int E, const std::vector<bool>& filter, bool predSet, bool copred)
  {
    E_x = E;
    E_dt = g.E_dt(E);
    E_extras = g.E_extras(E);
    E_lagged_extras = g.num_extras_lagged * E;
    E_actual = g.E_actual(E);
    bool all = filter.size() == 0;
    for (int i = 0; i < (int)g.t.size(); i++) {
      if (!(all || filter[i]))
        continue;
      int ti = i;
      double target = g.get_target(i, copred, predSet, ti);
      if (!predSet && target == MISSING_D)
        continue; // src/manifold.h:185-190
      rows.push_back(i);
      pointTimes.push_back(g.t[i]);
      targets.push_back(target);
      targetTimes.push_back(target != MISSING_D ? g.t[ti] : MISSING_D);
      if (g.panel_mode)
        panelIDs.push_back(g.panelIDs[i]);
    }
    n = (int)rows.size();
    flat.resize((size_t)n * E_actual);
    for (int r = 0; r < n; r++)
      g.fill_in_point(rows[r], E_x, copred, predSet, g.dtWeight, &flat[(size_t)r * E_actual]);
  }

  orc_manifold_view view() const
  {
    orc_manifold_view v;
    v.data = flat.data();
    v.targets = targets.data();
    v.point_times = pointTimes.data();
    v.target_times = targetTimes.data();
    v.panel = panelIDs.empty() ? nullptr : panelIDs.data();
    v.num_points = n;
    v.E_x = E_x;
    v.E_dt = E_dt;
    v.E_extras = E_extras;
    v.E_lagged_extras = E_lagged_extras;
    v.E_actual = E_actual;
    return v;
  }
};

/* ---------------------------------------------------------------- distances (src/distances.cpp) */

struct Pairs
{
  std::vector<int> inds;
  std::vector<double> dists;
};

static inline double id_weight(const orc_options& o, int pl, int pp)
{
  return (double)o.id_weights[(size_t)pl * o.num_panels + pp];
}

// src/distances.cpp:151-229 (eager) and :69-149 (lazy: no future-point skip, missing and
// categorical always checked).  MAE uses |.| (see the header note on the reference's unqualified abs).
static Pairs lp_distances(int Mp_i, const orc_options& o, const orc_manifold_view& M, const orc_manifold_view& Mp)
{
  Pairs out;
  out.inds.reserve(M.num_points);
  out.dists.reserve(M.num_points);
  const int E = M.E_actual;
  const bool lazy = o.low_memory != 0;
  const bool hasMissing = lazy || o.has_missing;
  const bool hasCat = lazy || o.has_categorical;
  const bool skipOther = o.panel_mode && o.idw < 0;
  const bool skipFuture = !lazy && o.use_only_past;
  const bool mae = o.distance == 0;
  const double* y = Mp.data + (size_t)Mp_i * E;
  for (int i = 0; i < M.num_points; i++) {
    if (skipOther && M.panel[i] != Mp.panel[Mp_i])
      continue;
    if (skipFuture && M.target_times[i] > Mp.point_times[Mp_i])
      continue;
    double acc = 0.0;
    if (o.panel_mode && o.idw > 0)
      acc += o.idw * (M.panel[i] != Mp.panel[Mp_i]);
    else if (o.panel_mode && o.id_weights != nullptr)
      acc += id_weight(o, M.panel[i], Mp.panel[Mp_i]);
    const double* x = M.data + (size_t)i * E;
    for (int j = 0; j < E; j++) {
      double d;
      if (hasMissing && (x[j] == MISSING_D || y[j] == MISSING_D)) {
        if (o.missing_distance == 0) {
          acc = MISSING_D;
          break;
        }
        d = o.missing_distance;
      } else if (hasCat && o.metrics != nullptr && o.metrics[j] == 1) {
        d = (x[j] != y[j]);
      } else {
        d = x[j] - y[j];
      }
      if (mae)
        acc += std::fabs(d) / E;
      else
        acc += d * d;
    }
    if (acc != 0 && acc != MISSING_D) {
      out.dists.push_back(mae ? acc : std::sqrt(acc));
      out.inds.push_back(i);
    }
  }
  return out;
}

/* Wasserstein distance (src/distances.cpp:296-527) is provided in wasserstein.cpp. */
Pairs wasserstein_distances(int Mp_i, const orc_options& o, const orc_manifold_view& M, const orc_manifold_view& Mp);

/* ---------------------------------------------------------------- k-NN (src/edm.cpp:652-681) */

// The first k of the ascending (distance, position) order.
static Pairs k_nearest(const Pairs& pot, int k)
{
  std::vector<int> idx(pot.inds.size());
  std::iota(idx.begin(), idx.end(), 0);
  auto cmp = [&](int a, int b) {
    if (pot.dists[a] != pot.dists[b])
      return pot.dists[a] < pot.dists[b];
    return a < b;
  };
  std::partial_sort(idx.begin(), idx.begin() + k, idx.end(), cmp);
  Pairs r;
  r.inds.resize(k);
  r.dists.resize(k);
  for (int i = 0; i < k; i++) {
    r.inds[i] = pot.inds[idx[i]];
    r.dists[i] = pot.dists[idx[i]];
  }
  return r;
}

/* ---------------------------------------------------------------- Jacobi SVD (Eigen 3.4 JacobiSVD) */

struct Rot
{
  double c, s;
};

// Rotation in the plane of rows p,q (Eigen applyOnTheLeft): x' = c x + s y, y' = -s x + c y.
static inline void rot_rows(double* A, int n, int p, int q, Rot j)
{
  if (j.c == 1.0 && j.s == 0.0)
    return;
  for (int i = 0; i < n; i++) {
    double xi = A[p * n + i], yi = A[q * n + i];
    A[p * n + i] = j.c * xi + j.s * yi;
    A[q * n + i] = -j.s * xi + j.c * yi;
  }
}

// Rotation of columns p,q by the transpose of j (Eigen applyOnTheRight).
static inline void rot_cols(double* A, int n, int p, int q, Rot j)
{
  Rot t{ j.c, -j.s };
  if (t.c == 1.0 && t.s == 0.0)
    return;
  for (int i = 0; i < n; i++) {
    double xi = A[i * n + p], yi = A[i * n + q];
    A[i * n + p] = t.c * xi + t.s * yi;
    A[i * n + q] = -t.s * xi + t.c * yi;
  }
}

// Eigen JacobiRotation::makeJacobi(x, y, z).
static inline Rot make_jacobi(double x, double y, double z)
{
  double deno = 2.0 * std::fabs(y);
  if (deno < DBL_MIN)
    return { 1.0, 0.0 };
  double tau = (x - z) / deno;
  double w = std::sqrt(tau * tau + 1.0);
  double t = tau > 0.0 ? 1.0 / (tau + w) : 1.0 / (tau - w);
  double sign_t = t > 0.0 ? 1.0 : -1.0;
  double nrm = 1.0 / std::sqrt(t * t + 1.0);
  double s = -sign_t * (y / std::fabs(y)) * std::fabs(t) * nrm;
  return { nrm, s };
}

// Eigen internal::real_2x2_jacobi_svd.
static inline void real_2x2(const double* A, int n, int p, int q, Rot* jl, Rot* jr)
{
  double m00 = A[p * n + p], m01 = A[p * n + q], m10 = A[q * n + p], m11 = A[q * n + q];
  Rot r1;
  double t = m00 + m11;
  double d = m10 - m01;
  if (std::fabs(d) < DBL_MIN) {
    r1 = { 1.0, 0.0 };
  } else {
    double u = t / d;
    double tmp = std::sqrt(1.0 + u * u);
    r1 = { u / tmp, 1.0 / tmp };
  }
  if (!(r1.c == 1.0 && r1.s == 0.0)) {
    double a0 = m00, a1 = m01, b0 = m10, b1 = m11;
    m00 = r1.c * a0 + r1.s * b0;
    m01 = r1.c * a1 + r1.s * b1;
    m10 = -r1.s * a0 + r1.c * b0;
    m11 = -r1.s * a1 + r1.c * b1;
  }
  (void)m10;
  *jr = make_jacobi(m00, m01, m11);
  Rot jt{ jr->c, -jr->s };
  jl->c = r1.c * jt.c - r1.s * jt.s;
  jl->s = r1.c * jt.s + r1.s * jt.c;
}

// JacobiSVD(A, ComputeThinU|ComputeThinV).solve(b) for square A (no QR preconditioner).
static void jacobi_solve(int n, const double* Ain, const double* b, double* x, double* sv_out = nullptr,
                         int* rank_out = nullptr)
{
  std::vector<double> W(Ain, Ain + (size_t)n * n), U((size_t)n * n, 0.0), V((size_t)n * n, 0.0), sv(n);
  double scale = 0.0;
  for (int i = 0; i < n * n; i++)
    scale = std::max(scale, std::fabs(Ain[i]));
  if (!std::isfinite(scale)) {
    for (int i = 0; i < n; i++)
      x[i] = 0.0; // Eigen reports InvalidInput; rank 0 solve yields zeros
    if (rank_out)
      *rank_out = 0;
    return;
  }
  if (scale == 0.0)
    scale = 1.0;
  for (int i = 0; i < n * n; i++)
    W[i] = Ain[i] / scale;
  for (int i = 0; i < n; i++) {
    U[i * n + i] = 1.0;
    V[i * n + i] = 1.0;
  }
  const double precision = 2.0 * DBL_EPSILON;
  const double considerAsZero = DBL_MIN;
  double maxDiag = 0.0;
  for (int i = 0; i < n; i++)
    maxDiag = std::max(maxDiag, std::fabs(W[i * n + i]));
  bool finished = false;
  while (!finished) {
    finished = true;
    for (int p = 1; p < n; p++) {
      for (int q = 0; q < p; q++) {
        double thr = std::max(considerAsZero, precision * maxDiag);
        if (std::fabs(W[p * n + q]) > thr || std::fabs(W[q * n + p]) > thr) {
          finished = false;
          Rot jl, jr;
          real_2x2(W.data(), n, p, q, &jl, &jr);
          rot_rows(W.data(), n, p, q, jl);
          rot_cols(U.data(), n, p, q, Rot{ jl.c, -jl.s });
          rot_cols(W.data(), n, p, q, jr);
          rot_cols(V.data(), n, p, q, jr);
          maxDiag = std::max(maxDiag, std::max(std::fabs(W[p * n + p]), std::fabs(W[q * n + q])));
        }
      }
    }
  }
  for (int i = 0; i < n; i++) {
    double a = W[i * n + i];
    sv[i] = std::fabs(a);
    if (a < 0.0)
      for (int r = 0; r < n; r++)
        U[r * n + i] = -U[r * n + i];
  }
  for (int i = 0; i < n; i++)
    sv[i] *= scale;
  int nonzero = n;
  for (int i = 0; i < n; i++) {
    int pos = i;
    double mx = sv[i];
    for (int j = i + 1; j < n; j++)
      if (sv[j] > mx) {
        mx = sv[j];
        pos = j;
      }
    if (mx == 0.0) {
      nonzero = i;
      break;
    }
    if (pos != i) {
      std::swap(sv[i], sv[pos]);
      for (int r = 0; r < n; r++) {
        std::swap(U[r * n + i], U[r * n + pos]);
        std::swap(V[r * n + i], V[r * n + pos]);
      }
    }
  }
  double thr = std::max(sv[0] * ((double)n * DBL_EPSILON), DBL_MIN);
  int r = nonzero - 1;
  while (r >= 0 && sv[r] < thr)
    r--;
  int rank = r + 1;
  std::vector<double> tmp(rank);
  for (int i = 0; i < rank; i++) {
    double s = 0.0;
    for (int j = 0; j < n; j++)
      s += U[j * n + i] * b[j];
    tmp[i] = (1.0 / sv[i]) * s;
  }
  for (int j = 0; j < n; j++) {
    double s = 0.0;
    for (int i = 0; i < rank; i++)
      s += V[j * n + i] * tmp[i];
    x[j] = s;
  }
  if (sv_out)
    for (int i = 0; i < n; i++)
      sv_out[i] = sv[i];
  if (rank_out)
    *rank_out = rank;
}


/* ---------------------------------------------------------------- weights' exp */

// exp(x) by FDLIBM's published method (Sun e_exp.c): x = k ln2 + r with a two-part ln2, a degree-5
// rational-polynomial exp(r), exact scaling by 2^k; within 1 ulp of exp.  The reference evaluates its
// simplex weights with libm exp (src/edm.cpp:722) and its S-map weights with Eigen's array exp
// (src/edm.cpp:745); both round within 1 ulp too, so the last bit of the weights is not pinned by the
// reference.  The oracle and the HIP kernels evaluate this same expression, so their FP64 weights agree
// bit for bit.
static double weight_exp(double x)
{
  const double ln2HI = 6.93147180369123816490e-01, ln2LO = 1.90821492927058770002e-10,
               invln2 = 1.44269504088896338700e+00;
  const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03, P3 = 6.61375632143793436117e-05,
               P4 = -1.65339022054652515390e-06, P5 = 4.13813679705723846039e-08;
  if (x != x)
    return x;
  if (x > 709.782712893383973096)
    return HUGE_VAL;
  if (x < -745.13321910194110842)
    return 0.0;
  const double ax = std::fabs(x);
  if (ax < 3.725290298461914e-09)
    return 1.0 + x;
  int k;
  double hi, lo;
  if (ax > 0.5 * 0.6931471805599453) {
    if (ax < 1.5 * 0.6931471805599453) {
      k = x < 0 ? -1 : 1;
      hi = x - k * ln2HI;
      lo = k * ln2LO;
    } else {
      k = (int)(invln2 * x + (x < 0 ? -0.5 : 0.5));
      const double t = k;
      hi = x - t * ln2HI;
      lo = t * ln2LO;
    }
  } else {
    k = 0;
    hi = x;
    lo = 0.0;
  }
  const double r = hi - lo;
  const double t = r * r;
  const double c = r - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
  if (k == 0)
    return 1.0 - ((r * c) / (c - 2.0) - r);
  const double y = 1.0 - ((lo - (r * c) / (2.0 - c)) - hi);
  return std::ldexp(y, k);
}

/* ---------------------------------------------------------------- predictors (src/edm.cpp:565-814) */

struct TaskOut
{
  double* pred;
  int* rc;
  double* coeffs;
  int* k_used;
  int* nn_idx;
  int nn_cap;
  int Np;
};

// src/edm.cpp:707-735
static void simplex(int Mp_i, int t, const orc_options& o, const orc_manifold_view& M, const Pairs& nn, TaskOut& out)
{
  int k = (int)nn.inds.size();
  double minDist = *std::min_element(nn.dists.begin(), nn.dists.end());
  std::vector<double> w(k);
  double sumw = 0.0;
  const double theta = o.thetas[t];
  for (int j = 0; j < k; j++) {
    w[j] = weight_exp(-theta * (nn.dists[j] / minDist));
    sumw = sumw + w[j];
  }
  double r = 0.0;
  for (int j = 0; j < k; j++)
    r = r + M.targets[nn.inds[j]] * (w[j] / sumw);
  out.pred[(size_t)t * out.Np + Mp_i] = r;
  out.rc[(size_t)t * out.Np + Mp_i] = SUCCESS;
}

// src/edm.cpp:737-814
static void smap(int Mp_i, int t, const orc_options& o, const orc_manifold_view& M, const orc_manifold_view& Mp,
                 const Pairs& nn, TaskOut& out)
{
  const int k = (int)nn.inds.size(), E = M.E_actual, n = E + 1;
  double mean = mean_seq(nn.dists.data(), k);
  std::vector<double> w(k), X((size_t)k * n), y(k);
  for (int i = 0; i < k; i++)
    w[i] = weight_exp(-o.thetas[t] * (nn.dists[i] / mean));
  for (int i = 0; i < k; i++) {
    const double* row = M.data + (size_t)nn.inds[i] * E;
    X[(size_t)i * n] = w[i];
    for (int j = 1; j < n; j++)
      X[(size_t)i * n + j] = w[i] * row[j - 1];
    y[i] = w[i] * M.targets[nn.inds[i]];
  }
  std::vector<double> A((size_t)n * n), b(n), ics(n);
  for (int p = 0; p < n; p++) {
    for (int q = p; q < n; q++) {
      double s = 0.0;
      for (int i = 0; i < k; i++)
        s += X[(size_t)i * n + p] * X[(size_t)i * n + q];
      A[p * n + q] = s;
      A[q * n + p] = s;
    }
    double s = 0.0;
    for (int i = 0; i < k; i++)
      s += X[(size_t)i * n + p] * y[i];
    b[p] = s;
  }
  jacobi_solve(n, A.data(), b.data(), ics.data());
  double r = ics[0];
  const double* yp = Mp.data + (size_t)Mp_i * E;
  for (int j = 0; j < E; j++)
    if (yp[j] != MISSING_D)
      r += yp[j] * ics[j + 1];
  if (o.save_smap_coeffs && out.coeffs && t == o.num_thetas - 1) {
    for (int j = 0; j < n; j++)
      out.coeffs[(size_t)Mp_i * n + j] = std::fabs(ics[j]) < 1.0e-11 ? MISSING_D : ics[j];
  }
  out.pred[(size_t)t * out.Np + Mp_i] = r;
  out.rc[(size_t)t * out.Np + Mp_i] = SUCCESS;
}

// src/edm.cpp:565-639
static void make_prediction(int Mp_i, const orc_options& o, const orc_manifold_view& M, const orc_manifold_view& Mp,
                            TaskOut& out)
{
  Pairs pot = o.distance == 2 ? wasserstein_distances(Mp_i, o, M, Mp) : lp_distances(Mp_i, o, M, Mp);
  int nValid = (int)pot.inds.size();
  int k = o.k;
  if (k > nValid) {
    if (o.force_compute) {
      k = nValid;
    } else {
      out.rc[Mp_i] = INSUFFICIENT_UNIQUE;
      return;
    }
  }
  if (k == 0 || nValid == 0) {
    out.rc[Mp_i] = o.force_compute ? SUCCESS : INSUFFICIENT_UNIQUE;
    return;
  }
  Pairs nn = (k < 0 || k == nValid) ? pot : k_nearest(pot, k);
  out.k_used[Mp_i] = (int)nn.inds.size();
  if (out.nn_idx) {
    int* dst = out.nn_idx + (size_t)Mp_i * out.nn_cap;
    for (int j = 0; j < out.nn_cap; j++)
      dst[j] = j < (int)nn.inds.size() ? nn.inds[j] : -1;
  }
  if (o.algorithm == 0) {
    for (int t = 0; t < o.num_thetas; t++)
      simplex(Mp_i, t, o, M, nn, out);
  } else if (o.algorithm == 1) {
    for (int t = 0; t < o.num_thetas; t++)
      smap(Mp_i, t, o, M, Mp, nn, out);
  } else {
    out.rc[Mp_i] = INVALID_ALGORITHM;
  }
}

// The edm_task output contract and fan-out (src/edm.cpp:338-432): outputs pre-filled, then one
// make_prediction per target on a pool of o.nthreads workers.
static void predict_all(const orc_manifold_view& M, const orc_manifold_view& Mp, const orc_options& o, TaskOut& out)
{
  const int Np = Mp.num_points, T = o.num_thetas, C = M.E_actual + 1;
  std::fill_n(out.pred, (size_t)T * Np, MISSING_D);
  std::fill_n(out.rc, (size_t)T * Np, (int)UNKNOWN_ERROR);
  if (out.coeffs)
    std::fill_n(out.coeffs, (size_t)Np * C, MISSING_D);
  std::fill_n(out.k_used, Np, MISSING_I);
  if (out.nn_idx)
    std::fill_n(out.nn_idx, (size_t)Np * out.nn_cap, -1);
  int nth = std::max(1, o.nthreads);
  if (nth <= 1 || Np < 2) {
    for (int i = 0; i < Np; i++)
      make_prediction(i, o, M, Mp, out);
    return;
  }
  std::atomic<int> next{ 0 };
  std::vector<std::thread> pool;
  for (int w = 0; w < nth; w++)
    pool.emplace_back([&] {
      for (;;) {
        int i = next.fetch_add(1);
        if (i >= Np)
          break;
        make_prediction(i, o, M, Mp, out);
      }
    });
  for (auto& th : pool)
    th.join();
}

/* ---------------------------------------------------------------- splitter (src/library_prediction_split.h) */

struct Splitter
{
  bool explore, full, shuffle;
  int crossfold, nUsable;
  std::vector<bool> usable, lib, pred;
  std::vector<int> group;
  Mt64 rng;

  // :28-48
  Splitter(bool e, bool f, bool s, int cf, const std::vector<bool>& u, int seed)
    : explore(e)
    , full(f)
    , shuffle(s)
    , crossfold(cf)
    , usable(u)
  {
    rng.init((uint64_t)(long long)seed);
    nUsable = (int)std::count(u.begin(), u.end(), true);
    if (cf > 0)
      setup_crossfold();
  }

  // :72-101
  void setup_crossfold()
  {
    std::vector<int> r;
    if (shuffle) {
      std::vector<double> u;
      for (int i = 0; i < nUsable; i++)
        u.push_back(rng.real2());
      r = rank_of(u);
    }
    int size = (int)std::round(((float)nUsable) / crossfold);
    int o = 0;
    group.assign(usable.size(), -1);
    for (size_t i = 0; i < usable.size(); i++) {
      if (!usable[i])
        continue;
      group[i] = shuffle ? r[o] % crossfold : o / size;
      o++;
    }
  }

  // :104-119
  int next_library_size(int iter) const
  {
    if (crossfold > 0) {
      int s = 0;
      for (int o = 0; o < nUsable; o++)
        if ((o + 1) % crossfold != (iter - 1))
          s++;
      return s;
    }
    return full ? nUsable : nUsable / 2;
  }

  // :145-157
  void update(int library, int iter)
  {
    size_t N = usable.size();
    lib.assign(N, false);
    pred.assign(N, false);
    if (explore && full) {
      lib = usable;
      pred = usable;
    } else if (explore && crossfold > 0) {
      for (size_t i = 0; i < N; i++)
        if (usable[i]) {
          bool inFold = group[i] == iter - 1;
          lib[i] = !inFold;
          pred[i] = inFold;
        }
    } else if (explore) {
      if (shuffle) {
        std::vector<double> u;
        for (int i = 0; i < nUsable; i++)
          u.push_back(rng.real2());
        double med = median(u);
        int o = 0;
        for (size_t i = 0; i < N; i++)
          if (usable[i]) {
            lib[i] = u[o] < med;
            pred[i] = !(u[o] < med);
            o++;
          }
      } else {
        int ls = nUsable / 2, o = 0;
        for (size_t i = 0; i < N; i++)
          if (usable[i]) {
            lib[i] = o < ls;
            pred[i] = !(o < ls);
            o++;
          }
      }
    } else {
      // :231-286 fixed_size_library
      pred = usable;
      if (shuffle) {
        std::vector<double> u;
        for (int i = 0; i < nUsable; i++)
          u.push_back(rng.real2());
        double cut = 1.0;
        if (library < (int)u.size()) {
          std::vector<double> c(u);
          std::nth_element(c.begin(), c.begin() + library, c.end());
          cut = c[library];
        }
        int o = 0;
        for (size_t i = 0; i < N; i++)
          if (usable[i]) {
            lib[i] = u[o] < cut;
            o++;
          }
      } else {
        int o = 0;
        for (size_t i = 0; i < N; i++)
          if (usable[i]) {
            lib[i] = o < library;
            o++;
          }
      }
    }
  }
};

/* ---------------------------------------------------------------- tasks (src/edm.cpp:90-215, 285-548) */

struct TaskSpec
{
  int E, k, library;
  bool copredict, savePrediction, saveSMAPCoeffs;
  std::vector<bool> lib, pred;
};

struct Stats
{
  int E, library;
  double theta, rho, mae;
};

struct Result
{
  int rc = 0;
  std::vector<Stats> stats, costats;
  int kMin = MISSING_I, kMax = MISSING_I;
  double missingdistance = 0, dtWeight = 0;
  bool hasPred = false, hasCopred = false, hasCoeffs = false;
  int predRows = 0, predCols = 0, copredRows = 0, copredCols = 0, coeffRows = 0, coeffCols = 0;
  std::vector<double> predM, copredM, coeffM;
  std::vector<std::vector<double>> Ms, Mps;
  std::vector<int> Mshape, Mpshape;
};

// src/r.cpp:73-105 — to_R_matrix column-major fill with a row filter.
static std::vector<double> to_R_matrix(const double* v, int r, int c, const std::vector<bool>& filter)
{
  std::vector<double> m((size_t)r * c);
  size_t o = 0;
  for (int j = 0; j < c; j++)
    for (int i = 0; i < r; i++) {
      if (!filter.empty() && !filter[i]) {
        m[(size_t)j * r + i] = MISSING_D;
        continue;
      }
      m[(size_t)j * r + i] = v[o++];
    }
  return m;
}

static std::vector<TaskSpec> configure_tasks(const Generator& g, const orc_command& c, int kIn, bool explore,
                                             const std::vector<int>& Es, const std::vector<int>& libraries,
                                             const std::vector<bool>& usable, bool copredictMode)
{
  std::vector<TaskSpec> tasks;
  Splitter sp(explore, c.full != 0, c.shuffle != 0, c.crossfold, usable, c.seed);
  int numLibraries = explore ? 1 : (int)libraries.size();
  int numIters = c.num_reps > c.crossfold ? c.num_reps : c.crossfold;
  int maxE = Es.back();
  std::vector<bool> cousable;
  if (copredictMode)
    cousable = g.generate_usable(maxE, true);
  int librarySize = 0;
  bool newSplit = true;
  for (int iter = 1; iter <= numIters; iter++) {
    if (explore) {
      newSplit = true;
      librarySize = sp.next_library_size(iter);
    }
    for (size_t ei = 0; ei < Es.size(); ei++) {
      int E = Es[ei];
      for (int l = 0; l == 0 || l < (int)libraries.size(); l++) {
        if (!explore)
          newSplit = true;
        int library = explore ? librarySize : libraries[l];
        int kAdj;
        if (kIn > 0)
          kAdj = kIn;
        else if (kIn < 0)
          kAdj = -1;
        else {
          int d = g.E_actual(E) + 1 + (c.algorithm == 1);
          kAdj = d < library ? d : library;
        }
        bool last = (E == maxE) && (l + 1 == numLibraries);
        TaskSpec ts;
        ts.E = E;
        ts.k = kAdj;
        ts.library = library;
        ts.copredict = false;
        if (explore)
          ts.savePrediction = c.save_final_predictions && ((iter == c.num_reps) || (c.crossfold > 0)) && last;
        else
          ts.savePrediction = c.save_final_predictions && (iter == c.num_reps) && last;
        ts.saveSMAPCoeffs = c.save_smap_coeffs != 0;
        if (newSplit) {
          sp.update(library, iter);
          newSplit = false;
        }
        ts.lib = sp.lib;
        ts.pred = sp.pred;
        tasks.push_back(ts);
        if (copredictMode) {
          TaskSpec co = ts;
          co.copredict = true;
          if (explore)
            co.savePrediction = c.save_final_copredictions && ((iter == c.num_reps) || (c.crossfold > 0)) && last;
          else
            co.savePrediction = c.save_final_copredictions && (iter == c.num_reps) && last;
          co.saveSMAPCoeffs = false;
          co.pred = cousable;
          tasks.push_back(co);
        }
      }
    }
  }
  return tasks;
}

static std::vector<double> replace_nan(const double* v, int n)
{
  std::vector<double> o(v, v + n);
  for (double& a : o)
    if (!std::isnormal(a))
      a = MISSING_D;
  return o;
}

static std::vector<int> expand_metrics(const Generator& g, int E, int distance)
{
  std::vector<int> m;
  if (distance == 2) {
    m.push_back(0);
    if (g.E_dt(E) > 0)
      m.push_back(0);
    for (int k = 0; k < g.num_extras; k++)
      m.push_back(0);
  } else {
    for (int l = 0; l < E + g.E_dt(E); l++)
      m.push_back(0);
    for (int k = 0; k < g.num_extras; k++) {
      int lags = k < g.num_extras_lagged ? E : 1;
      for (int l = 0; l < lags; l++)
        m.push_back(0);
    }
  }
  return m;
}

static void build_generator(const orc_command& c, Generator& g, bool& explore)
{
  int n = c.n;
  g.t = replace_nan(c.t, n);
  g.x = replace_nan(c.x, n);
  for (int i = 0; i < n; i++)
    if (g.t[i] == MISSING_D)
      g.x[i] = MISSING_D; // src/r.cpp:179-183
  explore = c.y == nullptr;
  if (!explore)
    g.xmap = replace_nan(c.y, n);
  if (c.panel)
    g.panelIDs.assign(c.panel, c.panel + n);
  if (c.co_x)
    g.co_x = replace_nan(c.co_x, n);
  for (int e = 0; e < c.num_extras; e++)
    g.extras.push_back(replace_nan(c.extras + (size_t)e * n, n));
  g.tau = c.tau;
  g.p = c.p;
  g.num_extras_lagged = std::max(0, std::min(c.num_extras_lagged, c.num_extras));
  g.allow_missing = c.allow_missing != 0;
  g.init(c.dt != 0, c.reldt != 0, c.dt_weight);
}

// src/r.cpp:107-468 — normalisation, task launch and result assembly.
static Result* run_command(const orc_command& c)
{
  Result* res = new Result();
  Generator g;
  bool explore;
  build_generator(c, g, explore);

  orc_options o;
  std::memset(&o, 0, sizeof(o));
  o.algorithm = c.algorithm;
  o.distance = c.distance;
  o.missing_distance = c.missing_distance;
  o.low_memory = c.low_memory;
  o.use_only_past = c.predict_with_past;
  o.force_compute = 1;
  o.save_smap_coeffs = c.save_smap_coeffs;
  o.nthreads = c.num_threads;
  o.aspect_ratio = c.aspect_ratio;
  std::vector<double> thetas = c.num_thetas > 0 ? std::vector<double>(c.thetas, c.thetas + c.num_thetas)
                                                : std::vector<double>{ 1.0 };
  o.num_thetas = (int)thetas.size();
  o.thetas = thetas.data();
  o.idw = c.panel_weight;
  o.panel_mode = c.panel != nullptr;

  // Panel weight matrix keyed like src/r.cpp:204-220 (consecutive-unique panel ids, later keys win).
  std::vector<float> idw_dense;
  std::vector<int> panel_dense;
  if (o.panel_mode && c.panel_weights != nullptr) {
    std::vector<int> uniq;
    for (int i = 0; i < c.n; i++)
      if (uniq.empty() || uniq.back() != c.panel[i])
        uniq.push_back(c.panel[i]);
    std::map<int, int> code;
    for (int id : uniq)
      if (!code.count(id))
        code[id] = (int)code.size();
    int P = (int)code.size();
    idw_dense.assign((size_t)P * P, 0.0f);
    std::vector<char> seen((size_t)P * P, 0);
    for (size_t i = 0; i < uniq.size(); i++)
      for (size_t j = 0; j < uniq.size(); j++) {
        int a = code[uniq[i]], b = code[uniq[j]];
        double v = (i < (size_t)c.panel_weights_dim && j < (size_t)c.panel_weights_dim)
                     ? c.panel_weights[i * c.panel_weights_dim + j]
                     : 0.0;
        idw_dense[(size_t)a * P + b] = (float)v;
        seen[(size_t)a * P + b] = 1;
      }
    o.num_panels = P;
    o.id_weights = idw_dense.data();
    o.idw = 0;
    panel_dense.resize(c.n);
    for (int i = 0; i < c.n; i++)
      panel_dense[i] = code[c.panel[i]];
    g.panelIDs = panel_dense;
  }

  o.has_missing = c.allow_missing;
  o.has_categorical = 0;
  if (c.allow_missing && o.missing_distance == 0)
    o.missing_distance = default_missing_distance(g.x);

  std::vector<int> Es(c.E, c.E + c.num_E);
  int maxE = Es.back();
  std::vector<bool> usable = g.generate_usable(maxE, false);
  int numUsable = (int)std::count(usable.begin(), usable.end(), true);
  if (numUsable == 0) {
    res->rc = 8000;
    return res;
  }
  std::vector<int> libraries(c.libraries, c.libraries + c.num_libraries);
  if (!explore && libraries.empty())
    libraries = { numUsable };
  bool copredictMode = c.co_x != nullptr;

  std::vector<TaskSpec> tasks = configure_tasks(g, c, c.k, explore, Es, libraries, usable, copredictMode);

  bool first = true;
  for (auto& ts : tasks) {
    std::vector<int> metrics = expand_metrics(g, ts.E, c.distance);
    orc_options to = o;
    to.k = ts.k;
    to.metrics = metrics.data();
    Manifold M(g, ts.E, ts.lib, false, ts.copredict);
    Manifold Mp(g, ts.E, ts.pred, true, ts.copredict);
    orc_manifold_view Mv = M.view(), Mpv = Mp.view();
    int Np = Mp.n, T = to.num_thetas, C = M.E_actual + 1;
    std::vector<double> pred((size_t)T * Np), coeffs((size_t)Np * C);
    std::vector<int> rc((size_t)T * Np), kUsed(Np);
    TaskOut out{ pred.data(), rc.data(), coeffs.data(), kUsed.data(), nullptr, 0, Np };
    predict_all(Mv, Mpv, to, out);

    std::vector<Stats>& st = ts.copredict ? res->costats : res->stats;
    for (int t = 0; t < T; t++) {
      std::vector<double> y1, y2;
      for (int i = 0; i < Np; i++)
        if (Mp.targets[i] != MISSING_D && pred[(size_t)t * Np + i] != MISSING_D) {
          y1.push_back(Mp.targets[i]);
          y2.push_back(pred[(size_t)t * Np + i]);
        }
      Stats s;
      s.library = ts.library;
      s.E = M.E_actual;
      s.theta = thetas[t];
      if (!y1.empty()) {
        s.mae = mean_absolute_error(y1, y2);
        s.rho = correlation(y1, y2);
      } else {
        s.mae = MISSING_D;
        s.rho = MISSING_D;
      }
      st.push_back(s);
    }
    int trc = *std::max_element(rc.begin(), rc.end());
    if (Np == 0)
      trc = 0;
    std::vector<int> ku;
    for (int v : kUsed)
      if (v != MISSING_I)
        ku.push_back(v);
    int kmin = MISSING_I, kmax = MISSING_I;
    if (!ku.empty()) {
      kmin = *std::min_element(ku.begin(), ku.end());
      kmax = *std::max_element(ku.begin(), ku.end());
    }
    if (first || kmin < res->kMin)
      res->kMin = kmin;
    if (first || kmax > res->kMax)
      res->kMax = kmax;
    first = false;
    res->rc = std::max(res->rc, trc);

    if (ts.savePrediction) {
      std::vector<double> last(Np);
      for (int i = 0; i < Np; i++)
        last[i] = pred[(size_t)(T - 1) * Np + i];
      auto m = to_R_matrix(last.data(), (int)ts.pred.size(), 1, ts.pred);
      if (!ts.copredict) {
        res->hasPred = true;
        res->predM = m;
        res->predRows = (int)ts.pred.size();
        res->predCols = 1;
      } else {
        res->hasCopred = true;
        res->copredM = m;
        res->copredRows = (int)ts.pred.size();
        res->copredCols = 1;
      }
    }
    if (ts.saveSMAPCoeffs) {
      res->hasCoeffs = true;
      res->coeffM = to_R_matrix(coeffs.data(), (int)ts.pred.size(), C, ts.pred);
      res->coeffRows = (int)ts.pred.size();
      res->coeffCols = C;
    }
    if (c.save_manifolds) {
      res->Ms.push_back(M.flat);
      res->Mps.push_back(Mp.flat);
      res->Mshape.push_back(M.n);
      res->Mshape.push_back(M.E_actual);
      res->Mpshape.push_back(Mp.n);
      res->Mpshape.push_back(Mp.E_actual);
    }
  }
  res->missingdistance = o.missing_distance;
  res->dtWeight = g.dtWeight;
  return res;
}

} // namespace orc

using namespace orc;

extern "C" {

int orc_predict(const orc_manifold_view* M, const orc_manifold_view* Mp, const orc_options* o, double* pred, int* rc,
                double* coeffs, int* k_used, int* nn_idx, int nn_cap)
{
  TaskOut out{ pred, rc, coeffs, k_used, nn_idx, nn_cap, Mp->num_points };
  predict_all(*M, *Mp, *o, out);
  return 0;
}

int orc_distances(const orc_manifold_view* M, const orc_manifold_view* Mp, const orc_options* o, int Mp_i, int* inds,
                  double* dists)
{
  Pairs pot = o->distance == 2 ? wasserstein_distances(Mp_i, *o, *M, *Mp) : lp_distances(Mp_i, *o, *M, *Mp);
  for (size_t a = 0; a < pot.inds.size(); a++) {
    inds[a] = pot.inds[a];
    dists[a] = pot.dists[a];
  }
  return (int)pot.inds.size();
}

void* orc_run_command(const orc_command* c)
{
  return run_command(*c);
}
int orc_result_rc(void* r)
{
  return ((Result*)r)->rc;
}
int orc_result_num_stats(void* r, int co)
{
  return (int)(co ? ((Result*)r)->costats.size() : ((Result*)r)->stats.size());
}
void orc_result_stats(void* r, int co, int* E, int* library, double* theta, double* rho, double* mae)
{
  auto& s = co ? ((Result*)r)->costats : ((Result*)r)->stats;
  for (size_t i = 0; i < s.size(); i++) {
    E[i] = s[i].E;
    library[i] = s[i].library;
    theta[i] = s[i].theta;
    rho[i] = s[i].rho;
    mae[i] = s[i].mae;
  }
}
void orc_result_k(void* r, int* kmin, int* kmax)
{
  *kmin = ((Result*)r)->kMin;
  *kmax = ((Result*)r)->kMax;
}
double orc_result_scalar(void* r, int which)
{
  return which == 0 ? ((Result*)r)->missingdistance : ((Result*)r)->dtWeight;
}
int orc_result_matrix_shape(void* r, int which, int* rows, int* cols)
{
  Result* R = (Result*)r;
  if (which == 0) {
    *rows = R->predRows;
    *cols = R->predCols;
    return R->hasPred;
  } else if (which == 1) {
    *rows = R->copredRows;
    *cols = R->copredCols;
    return R->hasCopred;
  }
  *rows = R->coeffRows;
  *cols = R->coeffCols;
  return R->hasCoeffs;
}
void orc_result_matrix(void* r, int which, double* out)
{
  Result* R = (Result*)r;
  auto& m = which == 0 ? R->predM : which == 1 ? R->copredM : R->coeffM;
  std::copy(m.begin(), m.end(), out);
}
int orc_result_num_manifolds(void* r)
{
  return (int)((Result*)r)->Ms.size();
}
void orc_result_manifold_shape(void* r, int i, int ps, int* rows, int* cols)
{
  auto& s = ps ? ((Result*)r)->Mpshape : ((Result*)r)->Mshape;
  *rows = s[2 * i];
  *cols = s[2 * i + 1];
}
void orc_result_manifold(void* r, int i, int ps, double* out)
{
  auto& m = ps ? ((Result*)r)->Mps[i] : ((Result*)r)->Ms[i];
  std::copy(m.begin(), m.end(), out);
}
void orc_result_free(void* r)
{
  delete (Result*)r;
}

void orc_mt64_stream(uint64_t seed, int n, double* out)
{
  Mt64 m;
  m.init(seed);
  for (int i = 0; i < n; i++)
    out[i] = m.real2();
}

int orc_split_masks(int n, const int* usable, int explore, int full, int shuffle, int crossfold, int seed, int num_reps,
                    int num_E, int num_libraries, const int* libraries, int* lib_out, int* pred_out, int max_tasks)
{
  std::vector<bool> u(n);
  for (int i = 0; i < n; i++)
    u[i] = usable[i] != 0;
  Splitter sp(explore != 0, full != 0, shuffle != 0, crossfold, u, seed);
  int numIters = num_reps > crossfold ? num_reps : crossfold;
  int task = 0;
  for (int iter = 1; iter <= numIters; iter++) {
    int ls = explore ? sp.next_library_size(iter) : 0;
    bool newSplit = true;
    for (int e = 0; e < num_E; e++)
      for (int l = 0; l == 0 || l < num_libraries; l++) {
        if (!explore)
          newSplit = true;
        int library = explore ? ls : libraries[l];
        if (newSplit) {
          sp.update(library, iter);
          newSplit = false;
        }
        if (task < max_tasks) {
          for (int i = 0; i < n; i++) {
            lib_out[(size_t)task * n + i] = sp.lib[i];
            pred_out[(size_t)task * n + i] = sp.pred[i];
          }
        }
        task++;
      }
  }
  return task;
}

int orc_usable(const orc_command* c, int maxE, int copred, int* usable_out)
{
  Generator g;
  bool explore;
  build_generator(*c, g, explore);
  auto u = g.generate_usable(maxE, copred != 0);
  int s = 0;
  for (size_t i = 0; i < u.size(); i++) {
    usable_out[i] = u[i];
    s += u[i];
  }
  return s;
}

int orc_generator_state(const orc_command* c, int maxE, double* t, double* x, double* xmap, double* co_x,
                        double* extras, int* obs, int* flags, double* dt_weight, int* usable_out)
{
  Generator g;
  bool explore;
  build_generator(*c, g, explore);
  const int n = c->n;
  for (int i = 0; i < n; i++) {
    t[i] = g.t[i];
    x[i] = g.x[i];
    obs[i] = g.obs[i];
    if (xmap)
      xmap[i] = g.xmap_mode ? g.xmap[i] : 0.0;
    if (co_x)
      co_x[i] = g.co_x.empty() ? 0.0 : g.co_x[i];
  }
  for (int e = 0; e < g.num_extras && extras; e++)
    for (int i = 0; i < n; i++)
      extras[(size_t)e * n + i] = g.extras[e][i];
  flags[0] = g.dt;
  flags[1] = g.reldt;
  flags[2] = g.xmap_mode;
  flags[3] = g.panel_mode;
  *dt_weight = g.dtWeight;
  auto u = g.generate_usable(maxE, false);
  int s = 0;
  for (int i = 0; i < n; i++) {
    usable_out[i] = u[i];
    s += u[i];
  }
  return s;
}

int orc_manifold(const orc_command* c, int E, const int* filter, int ps, int copred, double* out, double* targets,
                 int max_points)
{
  Generator g;
  bool explore;
  build_generator(*c, g, explore);
  std::vector<bool> f(c->n);
  for (int i = 0; i < c->n; i++)
    f[i] = filter[i] != 0;
  Manifold M(g, E, f, ps != 0, copred != 0);
  int n = std::min(M.n, max_points);
  std::copy(M.flat.begin(), M.flat.begin() + (size_t)n * M.E_actual, out);
  std::copy(M.targets.begin(), M.targets.begin() + n, targets);
  return M.n;
}

void orc_jacobi_solve(int n, const double* A, const double* b, double* x, double* sv, int* rank)
{
  jacobi_solve(n, A, b, x, sv, rank);
}

double orc_correlation(int n, const double* a, const double* b)
{
  return correlation(std::vector<double>(a, a + n), std::vector<double>(b, b + n));
}
double orc_mae(int n, const double* a, const double* b)
{
  return mean_absolute_error(std::vector<double>(a, a + n), std::vector<double>(b, b + n));
}
}

extern "C" double orc_weight_exp(double x)
{
  return orc::weight_exp(x);
}
