// Host side of the MI355X fastEDM prediction core: the reference's run_command (src/r.cpp:107-468)
// and its task machinery (src/edm.cpp:90-548), feeding the GPU through edm_predict.
//
// Layout choices made for the GPU rather than taken from the reference:
//  * lag resolution and point filling are split: every row's lag rows (and target row) are resolved
//    once per E into an index table (lag_rows), and a manifold is then a gather of that table, so a
//    task's library and prediction manifolds are row subsets of one per-E point table;
//  * tasks that share E and prediction rows are dispatched with one device context and reused
//    device buffers.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <string>
#include <sys/stat.h>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <random>
#include <vector>

#include "../../include/fastedm_hip.h"

namespace fedmh {

constexpr double MISS = FEDM_MISSING_D;
constexpr int MISS_I = FEDM_MISSING_I;

// ------------------------------------------------------------------ statistics (src/stats.cpp)
static double seq_mean(const std::vector<double>& v)
{
  double s = 0;
  for (double a : v)
    s += a;
  return s / (double)v.size();
}

static double sample_sd(const std::vector<double>& v) // src/stats.cpp:96-101
{
  const double m = seq_mean(v);
  double s = 0;
  for (double a : v)
    s += (a - m) * (a - m);
  return std::sqrt(s / (double)(v.size() - 1));
}

static std::vector<double> observed(const std::vector<double>& v)
{
  std::vector<double> o;
  o.reserve(v.size());
  for (double a : v)
    if (a != MISS)
      o.push_back(a);
  return o;
}

static double pearson(const std::vector<double>& a, const std::vector<double>& b) // src/stats.cpp:64-82
{
  const double ma = seq_mean(a), mb = seq_mean(b);
  double saa = 0, sbb = 0, sab = 0;
  for (size_t i = 0; i < a.size(); i++) {
    const double da = a[i] - ma, db = b[i] - mb;
    saa += da * da;
    sbb += db * db;
  }
  const double sa = std::sqrt(saa), sb = std::sqrt(sbb);
  if (sa < 1e-9 || sb < 1e-9)
    return MISS;
  for (size_t i = 0; i < a.size(); i++)
    sab += (a[i] - ma) * (b[i] - mb);
  return sab / (sa * sb);
}

static double mae_of(const std::vector<double>& a, const std::vector<double>& b) // src/stats.cpp:84-94
{
  double s = 0;
  for (size_t i = 0; i < a.size(); i++)
    s += std::fabs(a[i] - b[i]);
  const double m = s / (double)a.size();
  return m < 1e-8 ? 0 : m;
}

// ------------------------------------------------------------------ MT19937-64 stream (vendor/mersennetwister.h)
// The reference generator returns the tempered seed first (src/vendor/mersennetwister.h:131,139),
// then the standard MT19937-64 sequence; values are (u >> 11) * 2^-53.
class Mt64Stream
{
  std::mt19937_64 eng_;
  uint64_t seed_;
  bool pending_seed_;

public:
  explicit Mt64Stream(uint64_t seed = 0)
    : eng_(seed)
    , seed_(seed)
    , pending_seed_(true)
  {
  }
  double next()
  {
    uint64_t u;
    if (pending_seed_) {
      pending_seed_ = false;
      u = seed_;
      u ^= (u >> 29) & 0x5555555555555555ULL;
      u ^= (u << 17) & 0x71D67FFFEDA60000ULL;
      u ^= (u << 37) & 0xFFF7EEE000000000ULL;
      u ^= (u >> 43);
    } else {
      u = eng_();
    }
    return (double)(u >> 11) * (1.0 / 9007199254740992.0);
  }
  std::vector<double> draw(int n)
  {
    std::vector<double> v(n);
    for (auto& a : v)
      a = next();
    return v;
  }
};

// ------------------------------------------------------------------ series + lag resolution (src/manifold.*)
struct Series
{
  std::vector<double> t, x, xmap, co_x;
  std::vector<std::vector<double>> extras;
  std::vector<int> panel, obs;
  int tau = 1, p = 0;
  int n_lagged = 0; // the first n_lagged extras enter a point with E lags each (src/manifold.h:152)
  bool dt = false, reldt = false, allow_missing = false, has_panel = false, has_xmap = false;
  double w_dt = 0;

  int n() const { return (int)t.size(); }
  int E_dt(int E) const { return (dt || reldt) ? E : 0; }
  int num_extras() const { return (int)extras.size(); }
  int E_extras(int E) const { return num_extras() + n_lagged * (E - 1); } // extra columns of a point
  int E_lagged_extras(int E) const { return n_lagged * E; }
  int E_actual(int E) const { return E + E_dt(E) + E_extras(E); }

  static double fgcd(double a, double b) // src/manifold.cpp:7-18
  {
    while (true) {
      if (a < b)
        std::swap(a, b);
      if (std::fabs(b) < 0.001)
        return a;
      const double r = a - std::floor(a / b) * b;
      a = b;
      b = r;
    }
  }

  void number_observations() // src/manifold.cpp:20-80
  {
    obs.assign(n(), MISS_I);
    if (!(dt || reldt)) {
      double unit = -1;
      for (int i = 1; i < n(); i++) {
        const double d = t[i] - t[i - 1];
        if (d <= 0 || t[i] == MISS || t[i - 1] == MISS)
          continue;
        unit = unit < 0 ? d : fgcd(d, unit);
      }
      const double t0 = *std::min_element(t.begin(), t.end());
      for (int i = 0; i < n(); i++)
        if (t[i] != MISS)
          obs[i] = (int)std::round((t[i] - t0) / unit);
    } else {
      int c = 0;
      for (int i = 0; i < n(); i++)
        if (t[i] != MISS && (allow_missing || x[i] != MISS))
          obs[i] = c++;
    }
  }

  // src/manifold.cpp:82-123: walk from row r in steps of dir looking for observation `want`.
  bool seek(int want, int& r, int dir, int pan) const
  {
    for (; r >= 0 && r < n(); r += dir) {
      if (has_panel) {
        if (panel[r] == MISS_I)
          continue;
        if (panel[r] != pan)
          return false;
      }
      const int o = obs[r];
      if (o == MISS_I)
        continue;
      if (o == want)
        return true;
      if ((dir > 0 && o > want) || (dir < 0 && o < want))
        return false;
    }
    return false;
  }

  static int add_wrap(int a, long long b) { return (int)(uint32_t)((uint32_t)a + (uint32_t)(int64_t)b); }

  int target_row(int i) const // src/manifold.cpp:281-312 (row index of the target, or -1)
  {
    int r = i;
    if (p != 0 && !seek(add_wrap(obs[i], p), r, p > 0 ? 1 : -1, has_panel ? panel[i] : -1))
      return -1;
    return r;
  }

  // Lag rows of row i for lags 0..E-1 (-1 where a lag is absent), src/manifold.cpp:169-190.
  void lag_rows(int i, int E, int* out) const
  {
    const int pan = has_panel ? panel[i] : -1;
    int r = i - 1;
    out[0] = i;
    for (int j = 1; j < E; j++)
      out[j] = seek(add_wrap(obs[i], -(long long)j * tau), r, -1, pan) ? r : -1;
  }

  double dt_column(int i) const // src/manifold.cpp:240-279 (the dt series used for the default weight)
  {
    if (t[i] == MISS)
      return MISS;
    int r = i - 1;
    if (!seek(add_wrap(obs[i], -(long long)tau), r, -1, has_panel ? panel[i] : -1))
      return MISS;
    if (!reldt)
      return t[i] - t[r];
    const int tr = target_row(i);
    return tr >= 0 ? t[tr] - t[r] : MISS;
  }

  // One manifold point (src/manifold.cpp:156-238) from resolved lag rows.
  void point(int i, int E, const int* lags, bool co, double w, double* out) const
  {
    const int Edt = E_dt(E);
    double tPred = MISS;
    if (dt || reldt) {
      const int tr = target_row(i);
      tPred = tr >= 0 ? t[tr] : MISS;
    }
    int prev = i;
    for (int j = 0; j < E; j++) {
      const int r = lags[j];
      // a missing lag keeps the previous lag row for the next dt difference (src/manifold.cpp:188-236)
      const int lagged = r >= 0 ? r : prev;
      out[j] = r < 0 ? MISS : (co ? co_x[r] : x[r]);
      if (Edt) {
        if (r < 0) {
          out[E + j] = MISS;
        } else {
          const double tNow = t[r];
          if (j == 0 || reldt)
            out[E + j] = (tNow != MISS && tPred != MISS) ? w * (tPred - tNow) : MISS;
          else {
            const double tNext = t[prev];
            out[E + j] = (tNext != MISS && tNow != MISS) ? w * (tNext - tNow) : MISS;
          }
        }
      }
      // extras: the lagged ones at every lag, the others once (src/manifold.cpp:225-233)
      for (int e = 0; e < num_extras(); e++) {
        if (e < n_lagged)
          out[E + Edt + e * E + j] = r < 0 ? MISS : extras[e][r];
        else if (j == 0)
          out[E + Edt + n_lagged * E + (e - n_lagged)] = extras[e][i];
      }
      prev = lagged;
    }
  }

  double target_value(int i, bool co, int* row) const
  {
    const int r = target_row(i);
    if (row)
      *row = r;
    if (r < 0)
      return MISS;
    return co ? co_x[r] : (has_xmap ? xmap[r] : x[r]);
  }

  std::vector<bool> usable(int maxE, bool co) const // src/manifold.cpp:337-360
  {
    std::vector<bool> u(n(), false);
    const int Ea = E_actual(maxE);
    std::vector<int> lags(maxE);
    std::vector<double> pt(Ea);
    for (int i = 0; i < n(); i++) {
      if (t[i] == MISS || (has_panel && panel[i] == MISS_I))
        continue;
      lag_rows(i, maxE, lags.data());
      point(i, maxE, lags.data(), co, 1.0, pt.data());
      bool any = false, all = true;
      for (double v : pt) {
        any |= v != MISS;
        all &= v != MISS;
      }
      u[i] = allow_missing ? any : all;
    }
    return u;
  }
};

// A manifold gathered for a row filter (src/manifold.h:172-220).
struct Gathered
{
  std::vector<double> data, targets, ptimes, ttimes;
  std::vector<int> panel;
  int n = 0, E = 0, Edt = 0, Eex = 0, Elag = 0, Ea = 0;
  edm_manifold_view view() const
  {
    edm_manifold_view v;
    v.data = data.data();
    v.targets = targets.data();
    v.point_times = ptimes.data();
    v.target_times = ttimes.data();
    v.panel = panel.empty() ? nullptr : panel.data();
    v.num_points = n;
    v.E_x = E;
    v.E_dt = Edt;
    v.E_extras = Eex;
    v.E_lagged_extras = Elag;
    v.E_actual = Ea;
    return v;
  }
};

static Gathered gather(const Series& s, int E, const std::vector<bool>& filter, bool prediction_set, bool copred)
{
  Gathered g;
  g.E = E;
  g.Edt = s.E_dt(E);
  g.Eex = s.E_extras(E);
  g.Elag = s.E_lagged_extras(E);
  g.Ea = s.E_actual(E);
  const bool co = copred && prediction_set;
  std::vector<int> rows;
  for (int i = 0; i < s.n(); i++) {
    if (!filter[i])
      continue;
    int tr;
    const double target = s.target_value(i, co, &tr);
    if (!prediction_set && target == MISS)
      continue;
    rows.push_back(i);
    g.targets.push_back(target);
    g.ptimes.push_back(s.t[i]);
    g.ttimes.push_back(target != MISS ? s.t[tr] : MISS);
    if (s.has_panel)
      g.panel.push_back(s.panel[i]);
  }
  g.n = (int)rows.size();
  g.data.resize((size_t)g.n * g.Ea);
  std::vector<int> lags(E);
  for (int r = 0; r < g.n; r++) {
    s.lag_rows(rows[r], E, lags.data());
    s.point(rows[r], E, lags.data(), co, s.w_dt, &g.data[(size_t)r * g.Ea]);
  }
  return g;
}

// ------------------------------------------------------------------ library/prediction splits
// (src/library_prediction_split.h)
struct Splits
{
  bool explore, full, shuffle;
  int crossfold, nu;
  std::vector<bool> usable, lib, pred;
  std::vector<int> fold;
  Mt64Stream rng;

  Splits(bool e, bool f, bool sh, int cf, const std::vector<bool>& u, int seed)
    : explore(e)
    , full(f)
    , shuffle(sh)
    , crossfold(cf)
    , usable(u)
    , rng((uint64_t)(long long)seed)
  {
    nu = (int)std::count(u.begin(), u.end(), true);
    if (cf > 0) {
      std::vector<int> rank1;
      if (shuffle) {
        const auto uu = rng.draw(nu);
        std::vector<int> order(nu);
        for (int i = 0; i < nu; i++)
          order[i] = i;
        std::sort(order.begin(), order.end(), [&](int a, int b) { return uu[a] < uu[b] || (uu[a] == uu[b] && a < b); });
        rank1.resize(nu);
        for (int r = 0; r < nu; r++)
          rank1[order[r]] = r + 1;
      }
      const int per = (int)std::round(((float)nu) / cf);
      fold.assign(u.size(), -1);
      int o = 0;
      for (size_t i = 0; i < u.size(); i++)
        if (u[i]) {
          fold[i] = shuffle ? rank1[o] % cf : o / per;
          o++;
        }
    }
  }

  int explore_library_size(int iter) const
  {
    if (crossfold > 0) {
      int s = 0;
      for (int o = 0; o < nu; o++)
        s += ((o + 1) % crossfold != (iter - 1));
      return s;
    }
    return full ? nu : nu / 2;
  }

  static double kth_value(std::vector<double> v, int k)
  {
    std::nth_element(v.begin(), v.begin() + k, v.end());
    return v[k];
  }

  void draw_split(int library, int iter)
  {
    const size_t N = usable.size();
    lib.assign(N, false);
    pred.assign(N, false);
    if (explore && full) {
      lib = usable;
      pred = usable;
      return;
    }
    if (explore && crossfold > 0) {
      for (size_t i = 0; i < N; i++)
        if (usable[i]) {
          lib[i] = fold[i] != iter - 1;
          pred[i] = !lib[i];
        }
      return;
    }
    if (explore) {
      if (shuffle) {
        const auto u = rng.draw(nu);
        double med;
        if (nu % 2 == 0)
          med = (kth_value(u, nu / 2 - 1) + kth_value(u, nu / 2)) / 2;
        else
          med = kth_value(u, nu / 2);
        int o = 0;
        for (size_t i = 0; i < N; i++)
          if (usable[i]) {
            lib[i] = u[o] < med;
            pred[i] = !lib[i];
            o++;
          }
      } else {
        int o = 0;
        for (size_t i = 0; i < N; i++)
          if (usable[i]) {
            lib[i] = o < nu / 2;
            pred[i] = !lib[i];
            o++;
          }
      }
      return;
    }
    pred = usable;
    if (shuffle) {
      const auto u = rng.draw(nu);
      const double cut = library < nu ? kth_value(u, library) : 1.0;
      int o = 0;
      for (size_t i = 0; i < N; i++)
        if (usable[i])
          lib[i] = u[o++] < cut;
    } else {
      int o = 0;
      for (size_t i = 0; i < N; i++)
        if (usable[i])
          lib[i] = o++ < library;
    }
  }
};

} // namespace fedmh

using namespace fedmh;

struct fedm_result
{
  int rc = 0;
  struct Row
  {
    int E, library;
    double theta, rho, mae;
  };
  std::vector<Row> stats, costats;
  int kmin = MISS_I, kmax = MISS_I;
  double missingdistance = 0, dtWeight = 0;
  std::vector<double> mat[3];
  int rows[3] = { 0, 0, 0 }, cols[3] = { 0, 0, 0 };
  bool has[3] = { false, false, false };
  std::vector<Gathered> Ms, Mps;
};

static std::vector<double> to_r_matrix(const double* v, int r, int c, const std::vector<bool>& filter)
{
  std::vector<double> m((size_t)r * c);
  size_t o = 0;
  for (int j = 0; j < c; j++)
    for (int i = 0; i < r; i++)
      m[(size_t)j * r + i] = filter[i] ? v[o++] : MISS;
  return m;
}

static std::vector<double> normalised(const double* v, int n)
{
  std::vector<double> o(v, v + n);
  for (double& a : o)
    if (!std::isnormal(a))
      a = MISS; // src/r.cpp:64-71: NaN, infinities, subnormals and 0.0 are missing
  return o;
}

extern "C" {

} // extern "C"

// A command turned into device work: per task its row lists, grouped into batches that share
// (E, copredict) and therefore one point table.
struct fedm_plan
{
  edm_ctx* ctx = nullptr;
  Series s;
  edm_options o;
  std::vector<double> thetas;
  std::vector<float> idw_table;
  int rc0 = 0;     // 8000 when no row is usable (src/r.cpp:260-266)
  bool copred_mode = false;
  int save_manifolds = 0, save_smap_coeffs = 0;
  std::vector<int> extra_metrics; // per extra: 0 Diff, 1 CheckSame (empty: all Diff)
  int (*keep_going)(void) = nullptr; // polled between tasks and stages (src/edm.cpp:134, 377, 419)
  bool broken = false;               // the last run stopped because keep_going said so
  bool fetched = false;              // group outputs copied back since the last run
  struct Task
  {
    int E, k, library, mask; // mask: index into masks (prediction rows)
    bool co, savePred;
    std::vector<int> lib;    // library rows (target present)
    int group = -1, slot = -1;
  };
  std::vector<Task> tasks;
  std::vector<std::vector<bool>> pred_masks;
  std::vector<std::vector<int>> pred_rows; // per mask
  struct Group
  {
    int E;
    bool co;
    std::vector<int> members; // task indices
    std::vector<double> table, targets, ttimes, ptable, ptimes;
    std::vector<int> panel;
    std::vector<int64_t> out_begin;
    int64_t total_out = 0;
    edm_batch* batch = nullptr;
    bool taskwise = false; // one edm_predict per task (the task-by-task plain route)
    bool want_coeffs = false;
    std::vector<double> pred;
    std::vector<int> rc, ku;
    std::vector<double> coeffs; // per task [Np x C] (S-map coefficient route)
    // the batch's arguments, kept for the autotuner's rebuilds
    std::vector<int> lb, lc, lr, pb, pc, pr, kk, metrics;
    edm_options go;
    std::string tune_key;       // workload class in the tuning cache
  };
  bool graph = false;
  std::vector<Group> groups;
  ~fedm_plan()
  {
    for (auto& g : groups)
      if (g.batch)
        edm_batch_destroy(g.batch);
  }
};

// Arguments of the task grid (src/edm.cpp:90-215).
struct TaskGrid
{
  std::vector<int> Es, libraries;
  std::vector<bool> usable;
  int k, num_reps, crossfold;
  bool explore, full, shuffle, save_predictions, save_copredictions, copredict;
  int seed;
};
static int plan_tasks(edm_ctx* ctx, fedm_plan* P, const TaskGrid& grid);
static std::string tune_key(const fedm_plan* P, const fedm_plan::Group& g);
static bool tune_cache_lookup(const std::string& key, fedm_knobs* k);
static int create_group_batch(fedm_plan* P, fedm_plan::Group& g, const fedm_knobs* kn, edm_batch** out);

static int plan_command(edm_ctx* ctx, const fedm_command* c, fedm_plan* P)
{
  P->ctx = ctx;
  P->keep_going = c->keep_going;
  const int n = c->n;
  Series& s = P->s;
  s.t = normalised(c->t, n);
  s.x = normalised(c->x, n);
  for (int i = 0; i < n; i++)
    if (s.t[i] == MISS)
      s.x[i] = MISS;
  const bool explore = c->y == nullptr;
  if (!explore) {
    s.xmap = normalised(c->y, n);
    s.has_xmap = true;
  }
  if (c->co_x)
    s.co_x = normalised(c->co_x, n);
  for (int e = 0; e < c->num_extras; e++)
    s.extras.push_back(normalised(c->extras + (size_t)e * n, n));
  s.tau = c->tau;
  s.p = c->p;
  s.allow_missing = c->allow_missing != 0;

  edm_options& o = P->o;
  std::memset(&o, 0, sizeof(o));
  o.algorithm = c->algorithm;
  o.distance = c->distance;
  o.missing_distance = c->missing_distance;
  o.low_memory = c->low_memory;
  o.use_only_past = c->predict_with_past;
  o.force_compute = 1;
  o.save_smap_coeffs = c->save_smap_coeffs;
  o.aspect_ratio = c->aspect_ratio;
  o.precision = c->precision;
  P->thetas = c->num_thetas > 0 ? std::vector<double>(c->thetas, c->thetas + c->num_thetas) : std::vector<double>{ 1.0 };
  o.num_thetas = (int)P->thetas.size();
  o.thetas = P->thetas.data();
  o.idw = c->panel_weight;
  o.panel_mode = c->panel != nullptr;
  P->save_manifolds = c->save_manifolds;
  P->save_smap_coeffs = c->save_smap_coeffs;

  if (c->panel) {
    s.panel.assign(c->panel, c->panel + n);
    s.has_panel = true;
    if (c->panel_weights) {
      // keys are the run-length-distinct panel ids in row order; a later key overwrites
      // (src/r.cpp:204-220). Panels are renumbered densely so the table can be indexed.
      std::vector<int> runs;
      for (int i = 0; i < n; i++)
        if (runs.empty() || runs.back() != c->panel[i])
          runs.push_back(c->panel[i]);
      std::map<int, int> code;
      for (int id : runs)
        if (id != MISS_I && !code.count(id))
          code.emplace(id, (int)code.size());
      const int PP = (int)code.size();
      P->idw_table.assign((size_t)PP * PP, 0.0f);
      const int D = c->panel_weights_dim;
      for (size_t a = 0; a < runs.size(); a++)
        for (size_t b = 0; b < runs.size(); b++) {
          if (runs[a] == MISS_I || runs[b] == MISS_I)
            continue;
          const double w = ((int)a < D && (int)b < D) ? c->panel_weights[a * D + b] : 0.0;
          P->idw_table[(size_t)code[runs[a]] * PP + code[runs[b]]] = (float)w;
        }
      for (int i = 0; i < n; i++)
        if (s.panel[i] != MISS_I)
          s.panel[i] = code[s.panel[i]];
      o.num_panels = PP;
      o.id_weights = P->idw_table.data();
      o.idw = 0;
    }
  }

  // dt weight (src/manifold.h:85-127)
  s.dt = c->dt != 0;
  s.reldt = c->reldt != 0;
  s.w_dt = c->dt_weight;
  if ((s.dt || s.reldt) && c->dt_weight == 0.0) {
    const bool dt0 = s.dt, rel0 = s.reldt;
    s.dt = true;
    s.reldt = false;
    s.number_observations();
    std::vector<double> dts(n);
    for (int i = 0; i < n; i++)
      dts[i] = s.dt_column(i);
    const auto xo = observed(s.x), dto = observed(dts);
    const double sdx = sample_sd(xo), sddt = sample_sd(dto);
    const double w = sddt == 0.0 ? -1 : sdx / sddt;
    if (w > 0) {
      s.dt = dt0;
      s.reldt = rel0;
      s.w_dt = w;
    } else {
      s.dt = s.reldt = false;
      s.w_dt = 0;
    }
  }
  s.number_observations();

  o.has_missing = c->allow_missing;
  if (c->allow_missing && o.missing_distance == 0) {
    const auto xo = observed(s.x);
    o.missing_distance = 2 / std::sqrt(M_PI) * sample_sd(xo);
  }

  std::vector<int> Es(c->E, c->E + c->num_E);
  const int maxE = Es.back();
  const auto usable = s.usable(maxE, false);
  std::vector<int> libraries(c->libraries, c->libraries + c->num_libraries);
  const int nu = (int)std::count(usable.begin(), usable.end(), true);
  if (!explore && libraries.empty() && nu > 0)
    libraries = { nu };
  TaskGrid grid{ Es, libraries, usable, c->k, c->num_reps, c->crossfold, explore, c->full != 0, c->shuffle != 0,
                 c->save_final_predictions != 0, c->save_final_copredictions != 0, c->co_x != nullptr, c->seed };
  return plan_tasks(ctx, P, grid);
}

// The task grid of launch_tasks (src/edm.cpp:90-215) over a prepared series and options; tasks keep
// the reference's order for the results.  Groups of tasks sharing (E, copredict) get one point table
// and one batched launch.
static int plan_tasks(edm_ctx* ctx, fedm_plan* P, const TaskGrid& grid)
{
  Series& s = P->s;
  edm_options& o = P->o;
  const int n = s.n();
  const std::vector<int>& Es = grid.Es;
  const std::vector<int>& libraries = grid.libraries;
  const std::vector<bool>& usable = grid.usable;
  const bool explore = grid.explore;
  const int maxE = Es.back();
  const int nu = (int)std::count(usable.begin(), usable.end(), true);
  if (nu == 0) {
    P->rc0 = 8000;
    return 0;
  }
  P->copred_mode = grid.copredict;
  std::vector<bool> cousable;
  if (P->copred_mode)
    cousable = s.usable(maxE, true);

  Splits sp(explore, grid.full, grid.shuffle, grid.crossfold, usable, grid.seed);
  const int numLibraries = explore ? 1 : (int)libraries.size();
  const int iters = grid.num_reps > grid.crossfold ? grid.num_reps : grid.crossfold;
  auto mask_id = [&](const std::vector<bool>& m) {
    for (size_t q = P->pred_masks.size(); q-- > 0;)
      if (P->pred_masks[q] == m)
        return (int)q;
    P->pred_masks.push_back(m);
    std::vector<int> rows;
    for (int i = 0; i < n; i++)
      if (m[i])
        rows.push_back(i);
    P->pred_rows.push_back(std::move(rows));
    return (int)P->pred_masks.size() - 1;
  };
  for (int iter = 1; iter <= iters; iter++) {
    int libSize = 0;
    bool fresh = true;
    if (explore)
      libSize = sp.explore_library_size(iter);
    for (size_t ei = 0; ei < Es.size(); ei++) {
      const int E = Es[ei];
      for (int l = 0; l == 0 || l < (int)libraries.size(); l++) {
        if (!explore)
          fresh = true;
        const int library = explore ? libSize : libraries[l];
        int kAdj;
        if (grid.k > 0)
          kAdj = grid.k;
        else if (grid.k < 0)
          kAdj = -1;
        else
          kAdj = std::min(s.E_actual(E) + 1 + (o.algorithm == FEDM_SMAP), library);
        const bool last = (E == maxE) && (l + 1 == numLibraries);
        const bool finalRep = iter == grid.num_reps || (explore && grid.crossfold > 0);
        const bool savePred = grid.save_predictions && finalRep && last;
        const bool saveCo = grid.save_copredictions && finalRep && last;
        if (fresh) {
          sp.draw_split(library, iter);
          fresh = false;
        }
        for (int twin = 0; twin < (P->copred_mode ? 2 : 1); twin++) {
          fedm_plan::Task tk;
          tk.E = E;
          tk.k = kAdj;
          tk.library = library;
          tk.co = twin == 1;
          tk.savePred = tk.co ? saveCo : savePred;
          tk.mask = mask_id(tk.co ? cousable : sp.pred);
          for (int i = 0; i < n; i++)
            if (sp.lib[i] && s.target_value(i, false, nullptr) != MISS) // src/manifold.h:185-190
              tk.lib.push_back(i);
          P->tasks.push_back(std::move(tk));
        }
      }
    }
  }

  // one point table per (E, copredict): library side uses x, prediction side co_x in copredict twins
  for (size_t ti = 0; ti < P->tasks.size(); ti++) {
    auto& tk = P->tasks[ti];
    int gi = -1;
    for (size_t q = 0; q < P->groups.size(); q++)
      if (P->groups[q].E == tk.E && P->groups[q].co == tk.co)
        gi = (int)q;
    if (gi < 0) {
      fedm_plan::Group g;
      g.E = tk.E;
      g.co = tk.co;
      const int Ea = s.E_actual(tk.E);
      g.table.resize((size_t)n * Ea);
      g.ptable.resize((size_t)n * Ea);
      g.targets.resize(n);
      g.ttimes.resize(n);
      g.ptimes = s.t;
      std::vector<int> lags(tk.E);
      for (int i = 0; i < n; i++) {
        if (s.t[i] == MISS) {
          std::fill(g.table.begin() + (size_t)i * Ea, g.table.begin() + (size_t)(i + 1) * Ea, MISS);
          std::fill(g.ptable.begin() + (size_t)i * Ea, g.ptable.begin() + (size_t)(i + 1) * Ea, MISS);
          g.targets[i] = MISS;
          g.ttimes[i] = MISS;
          continue;
        }
        s.lag_rows(i, tk.E, lags.data());
        s.point(i, tk.E, lags.data(), false, s.w_dt, &g.table[(size_t)i * Ea]);
        if (tk.co)
          s.point(i, tk.E, lags.data(), true, s.w_dt, &g.ptable[(size_t)i * Ea]);
        int tr;
        g.targets[i] = s.target_value(i, false, &tr);
        g.ttimes[i] = g.targets[i] != MISS ? s.t[tr] : MISS;
      }
      if (!tk.co)
        g.ptable = g.table;
      if (s.has_panel)
        g.panel = s.panel;
      P->groups.push_back(std::move(g));
      gi = (int)P->groups.size() - 1;
    }
    auto& g = P->groups[gi];
    tk.group = gi;
    tk.slot = (int)g.members.size();
    g.members.push_back((int)ti);
  }
  const int T = o.num_thetas;
  for (auto& g : P->groups) {
    const int Ea = s.E_actual(g.E);
    std::vector<int>&lb = g.lb, &lc = g.lc, &lr = g.lr, &pb = g.pb, &pc = g.pc, &pr = g.pr, &kk = g.kk;
    std::vector<int> mask_begin(P->pred_rows.size(), -1);
    for (int ti : g.members) {
      const auto& tk = P->tasks[ti];
      lb.push_back((int)lr.size());
      lc.push_back((int)tk.lib.size());
      lr.insert(lr.end(), tk.lib.begin(), tk.lib.end());
      if (mask_begin[tk.mask] < 0) {
        mask_begin[tk.mask] = (int)pr.size();
        pr.insert(pr.end(), P->pred_rows[tk.mask].begin(), P->pred_rows[tk.mask].end());
      }
      pb.push_back(mask_begin[tk.mask]);
      pc.push_back((int)P->pred_rows[tk.mask].size());
      kk.push_back(tk.k);
      g.out_begin.push_back(g.total_out);
      g.total_out += (int64_t)T * P->pred_rows[tk.mask].size();
    }
    // expanded metrics (src/distances.cpp:29-67): x, dt and lagged coordinates Diff, each extra its own
    std::vector<int>& metrics = g.metrics;
    metrics.assign(Ea, 0);
    if (!P->extra_metrics.empty()) {
      if (o.distance == FEDM_DIST_WASSERSTEIN) {
        // one per variable: x, dt, then the extras
        const int rows = 1 + (s.E_dt(g.E) > 0);
        for (int e = 0; e < s.num_extras() && rows + e < Ea; e++)
          metrics[rows + e] = P->extra_metrics[e];
      } else {
        // one per coordinate: a lagged extra's metric at each of its E lags
        int c = g.E + s.E_dt(g.E);
        for (int e = 0; e < s.num_extras(); e++)
          for (int l = 0; l < (e < s.n_lagged ? g.E : 1); l++)
            metrics[c++] = P->extra_metrics[e];
      }
    }
    edm_options& go = g.go;
    go = o;
    go.metrics = metrics.data();
    const bool want_coeffs = !g.co && P->save_smap_coeffs && o.algorithm == FEDM_SMAP;
    go.save_smap_coeffs = want_coeffs;
    g.want_coeffs = want_coeffs;
    g.taskwise = o.precision == FEDM_PLAIN_TASKWISE;
    if (g.taskwise)
      continue;
    g.tune_key = tune_key(P, g);
    fedm_knobs kn;
    const bool cached = tune_cache_lookup(g.tune_key, &kn);
    const int r = create_group_batch(P, g, cached ? &kn : nullptr, &g.batch);
    if (r != 0)
      return r;
  }
  return 0;
}

static int create_group_batch(fedm_plan* P, fedm_plan::Group& g, const fedm_knobs* kn, edm_batch** out)
{
  const Series& s = P->s;
  const int n = s.n(), Ea = s.E_actual(g.E);
  edm_manifold_view Lv{ g.table.data(), g.targets.data(), g.ptimes.data(), g.ttimes.data(),
                        g.panel.empty() ? nullptr : g.panel.data(), n, g.E, s.E_dt(g.E), s.E_extras(g.E),
                        s.E_lagged_extras(g.E), Ea };
  edm_manifold_view Pv{ g.ptable.data(), g.targets.data(), g.ptimes.data(), g.ttimes.data(),
                        g.panel.empty() ? nullptr : g.panel.data(), n, g.E, s.E_dt(g.E), s.E_extras(g.E),
                        s.E_lagged_extras(g.E), Ea };
  return edm_batch_create_knobs(P->ctx, &Lv, &Pv, &g.go, (int)g.members.size(), g.lb.data(), g.lc.data(),
                                g.lr.data(), g.pb.data(), g.pc.data(), g.pr.data(), g.kk.data(), g.out_begin.data(),
                                kn, out);
}

// ------------------------------------------------------------------ autotuner and its cache

static int log2_bucket(long long v)
{
  int b = 0;
  while (v > 1) {
    v >>= 1;
    b++;
  }
  return b;
}

// The workload class of a group: what the best launch choices depend on, independent of the route the
// choices themselves select.
static std::string tune_key(const fedm_plan* P, const fedm_plan::Group& g)
{
  const edm_options& o = g.go;
  int kmax = 0, lmax = 0;
  for (size_t i = 0; i < g.kk.size(); i++) {
    kmax = std::max(kmax, g.kk[i]);
    lmax = std::max(lmax, g.lc[i]);
  }
  const bool generic = o.has_missing || o.has_categorical || o.panel_mode || o.use_only_past || o.low_memory ||
                       o.distance != FEDM_DIST_EUCLIDEAN;
  char buf[256];
  snprintf(buf, sizeof buf, "alg%d|prec%d|gen%d|dist%d|Ea%d|k%d|T%d|tasks%d|lib%d|pred%d|coef%d", o.algorithm,
           o.precision, generic ? 1 : 0, o.distance, P->s.E_actual(g.E), kmax, o.num_thetas,
           log2_bucket((long long)g.kk.size()), log2_bucket(lmax), log2_bucket((long long)g.pr.size()),
           g.want_coeffs ? 1 : 0);
  return buf;
}

static std::string tune_cache_path()
{
  if (const char* e = getenv("FEDM_TUNE_CACHE"))
    return e;
  const char* h = getenv("HOME");
  return std::string(h ? h : ".") + "/.cache/fastedm_hip/tune.txt";
}

static std::string knobs_text(const fedm_knobs& k)
{
  char buf[384];
  snprintf(buf, sizeof buf,
           "rank_scale=%.3f direct_blocks_per_cu=%d direct_alt_waves=%d list_blocks=%d mfma_blocks=%d use_mfma=%d "
           "smap_lds=%d direct_prune=%d concurrent_routes=%d smap_block=%d mfma_seed=%d mfma_prune=%d "
           "mfma_split=%d",
           k.rank_scale, k.direct_blocks_per_cu, k.direct_alt_waves, k.list_blocks, k.mfma_blocks, k.use_mfma,
           k.smap_lds, k.direct_prune, k.concurrent_routes, k.smap_block, k.mfma_seed, k.mfma_prune,
           k.mfma_split);
  return buf;
}

static bool parse_knobs(const std::string& line, fedm_knobs* k)
{
  fedm_default_knobs(k);
  int got = 0;
  size_t pos = 0;
  while (pos < line.size()) {
    const size_t end = line.find(' ', pos);
    const std::string tok = line.substr(pos, end == std::string::npos ? std::string::npos : end - pos);
    pos = end == std::string::npos ? line.size() : end + 1;
    const size_t eq = tok.find('=');
    if (eq == std::string::npos)
      continue;
    const std::string name = tok.substr(0, eq);
    const double v = atof(tok.c_str() + eq + 1);
    got++;
    if (name == "rank_scale")
      k->rank_scale = v;
    else if (name == "direct_blocks_per_cu")
      k->direct_blocks_per_cu = (int)v;
    else if (name == "direct_alt_waves")
      k->direct_alt_waves = (int)v;
    else if (name == "list_blocks")
      k->list_blocks = (int)v;
    else if (name == "mfma_blocks")
      k->mfma_blocks = (int)v;
    else if (name == "use_mfma")
      k->use_mfma = (int)v;
    else if (name == "smap_lds")
      k->smap_lds = (int)v;
    else if (name == "direct_prune")
      k->direct_prune = (int)v;
    else if (name == "concurrent_routes")
      k->concurrent_routes = (int)v;
    else if (name == "smap_block")
      k->smap_block = (int)v;
    else if (name == "mfma_seed")
      k->mfma_seed = (int)v;
    else if (name == "mfma_prune")
      k->mfma_prune = (int)v;
    else if (name == "mfma_split")
      k->mfma_split = (int)v;
    else
      got--;
  }
  return got > 0;
}

// The cache: one line per workload class, "<class>\t<knobs>\t<ms>"; a later line for a class wins.
static std::map<std::string, std::pair<std::string, std::string>> tune_cache_read()
{
  std::map<std::string, std::pair<std::string, std::string>> m;
  FILE* f = fopen(tune_cache_path().c_str(), "r");
  if (!f)
    return m;
  char buf[1024];
  while (fgets(buf, sizeof buf, f)) {
    std::string line(buf);
    while (!line.empty() && (line.back() == '\n' || line.back() == '\r'))
      line.pop_back();
    const size_t a = line.find('\t');
    if (a == std::string::npos)
      continue;
    const size_t b = line.find('\t', a + 1);
    m[line.substr(0, a)] = { line.substr(a + 1, b == std::string::npos ? std::string::npos : b - a - 1),
                             b == std::string::npos ? "" : line.substr(b + 1) };
  }
  fclose(f);
  return m;
}

static bool tune_cache_lookup(const std::string& key, fedm_knobs* k)
{
  const auto m = tune_cache_read();
  const auto it = m.find(key);
  return it != m.end() && parse_knobs(it->second.first, k);
}

static void tune_cache_write(const std::string& key, const fedm_knobs& k, double ms)
{
  auto m = tune_cache_read();
  char t[64];
  snprintf(t, sizeof t, "%.4f", ms);
  m[key] = { knobs_text(k), t };
  const std::string path = tune_cache_path();
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos) {
    // create the directory chain (best effort)
    std::string dir;
    for (size_t i = 0; i <= slash; i++) {
      if (path[i] == '/' && i > 0)
        mkdir(dir.c_str(), 0755);
      dir += path[i];
    }
  }
  FILE* f = fopen(path.c_str(), "w");
  if (!f)
    return;
  for (const auto& kv : m)
    fprintf(f, "%s\t%s\t%s\n", kv.first.c_str(), kv.second.first.c_str(), kv.second.second.c_str());
  fclose(f);
}

// Median wall time of `trials` runs after one warm-up run, in ms (negative on an error).
static double time_batch(edm_batch* b, int trials)
{
  if (edm_batch_run(b) != 0)
    return -1.0;
  std::vector<double> t;
  for (int i = 0; i < trials; i++) {
    const auto t0 = std::chrono::steady_clock::now();
    if (edm_batch_run(b) != 0)
      return -1.0;
    t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  }
  std::sort(t.begin(), t.end());
  return t[t.size() / 2];
}

struct BatchOutputs
{
  std::vector<double> pred;
  std::vector<int> rc, ku;
  bool operator==(const BatchOutputs& o) const
  {
    return pred.size() == o.pred.size() && std::memcmp(pred.data(), o.pred.data(), pred.size() * sizeof(double)) == 0 &&
           rc == o.rc && ku == o.ku;
  }
};

static int fetch_outputs(edm_batch* b, const fedm_plan::Group& g, int T, BatchOutputs& out)
{
  out.pred.assign((size_t)g.total_out, 0.0);
  out.rc.assign((size_t)g.total_out, 0);
  out.ku.assign((size_t)(g.total_out / T), 0);
  return edm_batch_fetch(b, out.pred.data(), out.rc.data(), out.ku.data());
}

static int tune_group(fedm_plan* P, fedm_plan::Group& g, int trials, std::string& report)
{
  const int T = P->o.num_thetas;
  const int route = edm_batch_route(g.batch);
  if (route == 0)
    return 0;
  fedm_knobs best;
  edm_batch_knobs(g.batch, &best);
  double best_ms = time_batch(g.batch, trials);
  if (best_ms < 0)
    return -3;
  BatchOutputs ref;
  int r = fetch_outputs(g.batch, g, T, ref);
  if (r != 0)
    return r;
  char line[512];
  snprintf(line, sizeof line, "group E=%d%s route=%s class=%s\n  start: %s  %.3f ms\n", g.E, g.co ? " copredict" : "",
           route == 2 ? "persistent" : "resident", g.tune_key.c_str(), knobs_text(best).c_str(), best_ms);
  report += line;
  // coordinate descent over the knobs of the route
  struct Axis
  {
    const char* name;
    std::vector<double> values;
  };
  std::vector<Axis> axes;
  const bool mixed = g.go.precision == FEDM_MIXED;
  const int Ea = P->s.E_actual(g.E);
  if (route == 2) {
    axes.push_back({ "direct_prune", { 1, 0 } });
    axes.push_back({ "concurrent_routes", { 1, 0 } });
    axes.push_back({ "rank_scale", { 1.0, 1.25, 1.5, 2.0, 3.0 } });
    axes.push_back({ "direct_alt_waves", { 0, 1 } });
    axes.push_back({ "direct_blocks_per_cu", { 0, 4, 2 } });
  } else {
    axes.push_back({ "list_blocks", { 512, 1024, 2048, 4096 } });
    if (mixed && Ea >= 8) {
      axes.push_back({ "use_mfma", { 1, 0 } });
      axes.push_back({ "mfma_seed", { 3, 2, 1, 0 } });
      axes.push_back({ "mfma_prune", { 1, 0 } });
      axes.push_back({ "mfma_split", { 0, 1, 4, 16, 32 } });
      axes.push_back({ "mfma_blocks", { 1024, 2048, 4096 } });
    }
    if (g.go.algorithm == FEDM_SMAP) {
      axes.push_back({ "smap_block", { 4, 0, 1, 2, 3 } });
      axes.push_back({ "smap_lds", { 0, 1 } });
    }
  }
  auto with = [](fedm_knobs k, const char* name, double v) {
    const std::string n(name);
    if (n == "rank_scale")
      k.rank_scale = v;
    else if (n == "direct_alt_waves")
      k.direct_alt_waves = (int)v;
    else if (n == "direct_blocks_per_cu")
      k.direct_blocks_per_cu = (int)v;
    else if (n == "list_blocks")
      k.list_blocks = (int)v;
    else if (n == "use_mfma")
      k.use_mfma = (int)v;
    else if (n == "mfma_blocks")
      k.mfma_blocks = (int)v;
    else if (n == "smap_lds")
      k.smap_lds = (int)v;
    else if (n == "direct_prune")
      k.direct_prune = (int)v;
    else if (n == "concurrent_routes")
      k.concurrent_routes = (int)v;
    else if (n == "smap_block")
      k.smap_block = (int)v;
    else if (n == "mfma_seed")
      k.mfma_seed = (int)v;
    else if (n == "mfma_prune")
      k.mfma_prune = (int)v;
    else if (n == "mfma_split")
      k.mfma_split = (int)v;
    return k;
  };
  for (const auto& ax : axes) {
    report += std::string("  ") + ax.name + ":";
    for (double v : ax.values) {
      const fedm_knobs cand = with(best, ax.name, v);
      if (std::memcmp(&cand, &best, sizeof cand) == 0) {
        snprintf(line, sizeof line, " %g %.3f ms (current) |", v, best_ms);
        report += line;
        continue;
      }
      edm_batch* nb = nullptr;
      r = create_group_batch(P, g, &cand, &nb);
      if (r != 0)
        return r;
      const double ms = time_batch(nb, trials);
      BatchOutputs got;
      const int fr = ms < 0 ? -3 : fetch_outputs(nb, g, T, got);
      if (fr != 0) {
        edm_batch_destroy(nb);
        return fr;
      }
      const bool same = got == ref;
      snprintf(line, sizeof line, " %g %.3f ms%s |", v, ms, same ? "" : " (outputs differ: rejected)");
      report += line;
      if (same && ms < best_ms) {
        edm_batch_destroy(g.batch);
        g.batch = nb;
        best = cand;
        best_ms = ms;
      } else {
        edm_batch_destroy(nb);
      }
    }
    report += "\n";
  }
  snprintf(line, sizeof line, "  best: %s  %.3f ms\n", knobs_text(best).c_str(), best_ms);
  report += line;
  if (P->graph)
    edm_batch_set_graph(g.batch, 1);
  tune_cache_write(g.tune_key, best, best_ms);
  return 1;
}

// One edm_predict per task of a group: the drop-in route of the reference's per-task seam.
static int run_group_taskwise(fedm_plan* P, fedm_plan::Group& g)
{
  const Series& s = P->s;
  const int T = P->o.num_thetas;
  const int Ea = s.E_actual(g.E), C = Ea + 1;
  g.pred.assign((size_t)g.total_out, MISS);
  g.rc.assign((size_t)g.total_out, FEDM_UNKNOWN_ERROR);
  g.ku.assign((size_t)(g.total_out / T), MISS_I);
  g.coeffs.assign((size_t)(g.total_out / T) * C, MISS);
  std::vector<int> metrics(Ea, 0);
  for (size_t m = 0; m < g.members.size(); m++) {
    const auto& tk = P->tasks[g.members[m]];
    std::vector<bool> lf(s.n(), false);
    for (int i : tk.lib)
      lf[i] = true;
    Gathered M = gather(s, tk.E, lf, false, tk.co);
    Gathered Mp = gather(s, tk.E, P->pred_masks[tk.mask], true, tk.co);
    edm_options to = P->o;
    if (to.precision == FEDM_PLAIN_TASKWISE)
      to.precision = FEDM_PLAIN;
    to.k = tk.k;
    to.metrics = metrics.data();
    const int64_t ob = g.out_begin[m];
    if (P->keep_going && !P->keep_going())
      return FEDM_BREAK_HIT;
    if (Mp.n > 0) {
      const edm_manifold_view mv = M.view(), mpv = Mp.view();
      const int r = edm_predict(P->ctx, &mv, &mpv, &to, g.pred.data() + ob, g.rc.data() + ob,
                                g.coeffs.data() + (size_t)(ob / T) * C, g.ku.data() + ob / T, nullptr, 0,
                                P->keep_going);
      if (r != 0)
        return r;
    }
  }
  return 0;
}

// Every group in turn; keep_going is polled before each group and passed on to the batches and the
// task-by-task route.  A stop returns FEDM_BREAK_HIT and leaves the plan marked as broken.
static int run_plan(fedm_plan* P)
{
  P->broken = false;
  P->fetched = false;
  for (auto& g : P->groups) {
    if (P->keep_going && !P->keep_going()) {
      P->broken = true;
      return FEDM_BREAK_HIT;
    }
    int r = 0;
    if (g.taskwise)
      r = run_group_taskwise(P, g);
    else if (g.batch) {
      edm_batch_set_keep_going(g.batch, P->keep_going);
      r = edm_batch_run(g.batch);
    }
    if (r == FEDM_BREAK_HIT)
      P->broken = true;
    if (r != 0)
      return r;
  }
  return 0;
}

// Copies the batched groups' outputs back to the host (the task-by-task groups hold theirs already).
static int fetch_groups(fedm_plan* P)
{
  if (P->fetched)
    return 0;
  const int T = P->o.num_thetas;
  for (auto& g : P->groups) {
    if (g.taskwise)
      continue; // filled by run_plan
    g.pred.assign((size_t)g.total_out, MISS);
    g.rc.assign((size_t)g.total_out, FEDM_UNKNOWN_ERROR);
    g.ku.assign((size_t)(g.total_out / T), MISS_I);
    if (g.batch) {
      int r = edm_batch_fetch(g.batch, g.pred.data(), g.rc.data(), g.ku.data());
      if (r == 0 && g.want_coeffs) {
        g.coeffs.assign((size_t)(g.total_out / T) * (P->s.E_actual(g.E) + 1), MISS);
        r = edm_batch_fetch_coeffs(g.batch, g.coeffs.data());
      }
      if (r != 0)
        return r;
    }
  }
  P->fetched = true;
  return 0;
}

static int finish_plan(fedm_plan* P, fedm_result* res)
{
  const Series& s = P->s;
  const int T = P->o.num_thetas;
  {
    const int r = fetch_groups(P);
    if (r != 0)
      return r;
  }
  bool first = true;
  for (auto& tk : P->tasks) {
    auto& g = P->groups[tk.group];
    const int64_t ob = g.out_begin[tk.slot];
    const auto& rows = P->pred_rows[tk.mask];
    const int Np = (int)rows.size(), Ea = s.E_actual(tk.E), C = Ea + 1;
    const double* pred = g.pred.data() + ob;
    const int* rc = g.rc.data() + ob;
    const int* ku = g.ku.data() + ob / T;
    std::vector<double> tg(Np);
    for (int i = 0; i < Np; i++)
      tg[i] = s.target_value(rows[i], tk.co, nullptr);
    auto& st = tk.co ? res->costats : res->stats;
    for (int th = 0; th < T; th++) {
      std::vector<double> a, b;
      for (int i = 0; i < Np; i++) {
        const double yp = pred[(size_t)th * Np + i];
        if (tg[i] != MISS && yp != MISS) {
          a.push_back(tg[i]);
          b.push_back(yp);
        }
      }
      fedm_result::Row r{ Ea, tk.library, P->thetas[th], MISS, MISS };
      if (!a.empty()) {
        r.mae = mae_of(a, b);
        r.rho = pearson(a, b);
      }
      st.push_back(r);
    }
    int trc = 0;
    for (int i = 0; i < T * Np; i++)
      trc = std::max(trc, rc[i]);
    int kmin = MISS_I, kmax = MISS_I;
    bool any = false;
    for (int i = 0; i < Np; i++)
      if (ku[i] != MISS_I) {
        kmin = any ? std::min(kmin, ku[i]) : ku[i];
        kmax = any ? std::max(kmax, ku[i]) : ku[i];
        any = true;
      }
    if (first || kmin < res->kmin)
      res->kmin = kmin;
    if (first || kmax > res->kmax)
      res->kmax = kmax;
    first = false;
    res->rc = std::max(res->rc, trc);
    const auto& mask = P->pred_masks[tk.mask];
    if (tk.savePred) {
      const int w = tk.co ? 1 : 0;
      res->mat[w] = to_r_matrix(pred + (size_t)(T - 1) * Np, (int)mask.size(), 1, mask);
      res->rows[w] = (int)mask.size();
      res->cols[w] = 1;
      res->has[w] = true;
    }
    if (!tk.co && P->save_smap_coeffs) {
      std::vector<double> cf((size_t)Np * C, MISS);
      if (!g.coeffs.empty())
        std::copy(g.coeffs.begin() + (size_t)(ob / T) * C, g.coeffs.begin() + (size_t)(ob / T + Np) * C, cf.begin());
      res->mat[2] = to_r_matrix(cf.data(), (int)mask.size(), C, mask);
      res->rows[2] = (int)mask.size();
      res->cols[2] = C;
      res->has[2] = true;
    }
    if (P->save_manifolds) {
      std::vector<bool> lf(s.n(), false);
      for (int i : tk.lib)
        lf[i] = true;
      res->Ms.push_back(gather(s, tk.E, lf, false, tk.co));
      res->Mps.push_back(gather(s, tk.E, mask, true, tk.co));
    }
  }
  res->missingdistance = P->o.missing_distance;
  res->dtWeight = s.w_dt;
  return 0;
}

extern "C" {

int fedm_plan_create(edm_ctx* ctx, const fedm_command* c, fedm_plan** out)
{
  *out = nullptr;
  auto P = std::make_unique<fedm_plan>();
  const int r = plan_command(ctx, c, P.get());
  if (r != 0)
    return r;
  *out = P.release();
  return 0;
}

int fedm_plan_run(fedm_plan* P)
{
  return run_plan(P);
}

int fedm_plan_result(fedm_plan* P, fedm_result** out)
{
  *out = nullptr;
  auto res = std::make_unique<fedm_result>();
  if (P->rc0 || P->broken) {
    // no usable rows, or an interrupted run: the reference stores no statistics (src/edm.cpp:438-439)
    res->rc = P->rc0 ? P->rc0 : FEDM_BREAK_HIT;
    *out = res.release();
    return 0;
  }
  const int r = finish_plan(P, res.get());
  if (r != 0)
    return r;
  *out = res.release();
  return 0;
}

int64_t fedm_plan_num_predictions(const fedm_plan* P)
{
  int64_t n = 0;
  for (const auto& g : P->groups)
    n += g.total_out;
  return n;
}

int fedm_plan_num_tasks(const fedm_plan* P)
{
  return (int)P->tasks.size();
}

int fedm_plan_tune(fedm_plan* P, int trials, char* report, int report_len)
{
  std::string rep;
  int tuned = 0;
  trials = std::max(trials, 1);
  for (auto& g : P->groups) {
    if (!g.batch)
      continue;
    const int r = tune_group(P, g, trials, rep);
    if (r < 0)
      return r;
    tuned += r;
  }
  P->fetched = false;
  if (report && report_len > 0) {
    const size_t n = std::min(rep.size(), (size_t)report_len - 1);
    std::memcpy(report, rep.data(), n);
    report[n] = 0;
  }
  return tuned;
}

int fedm_plan_group_knobs(const fedm_plan* P, int group, fedm_knobs* out)
{
  if (group < 0 || group >= (int)P->groups.size() || !P->groups[group].batch)
    return -22;
  edm_batch_knobs(P->groups[group].batch, out);
  return 0;
}

int fedm_plan_set_graph(fedm_plan* P, int on)
{
  P->graph = on != 0;
  int n = 0;
  for (auto& g : P->groups)
    if (g.batch)
      n += edm_batch_set_graph(g.batch, on);
  return n;
}

int fedm_plan_task(const fedm_plan* P, int task, int* info, int* lib_rows, int* pred_rows)
{
  if (task < 0 || task >= (int)P->tasks.size())
    return -22;
  const auto& tk = P->tasks[task];
  const auto& rows = P->pred_rows[tk.mask];
  info[0] = P->s.E_actual(tk.E);
  info[1] = tk.library;
  info[2] = tk.k;
  info[3] = (int)tk.lib.size();
  info[4] = (int)rows.size();
  info[5] = tk.co ? 1 : 0;
  info[6] = tk.E;
  if (lib_rows)
    std::copy(tk.lib.begin(), tk.lib.end(), lib_rows);
  if (pred_rows)
    std::copy(rows.begin(), rows.end(), pred_rows);
  return 0;
}

int fedm_plan_fetch(fedm_plan* P, double* pred, int* rc, int* k_used)
{
  if (P->rc0)
    return 0;
  const int r = fetch_groups(P);
  if (r != 0)
    return r;
  const int T = P->o.num_thetas;
  size_t o = 0, ko = 0;
  for (const auto& tk : P->tasks) {
    const auto& g = P->groups[tk.group];
    const int64_t ob = g.out_begin[tk.slot];
    const size_t Np = P->pred_rows[tk.mask].size();
    std::copy(g.pred.begin() + ob, g.pred.begin() + ob + (int64_t)(T * Np), pred + o);
    std::copy(g.rc.begin() + ob, g.rc.begin() + ob + (int64_t)(T * Np), rc + o);
    std::copy(g.ku.begin() + ob / T, g.ku.begin() + ob / T + (int64_t)Np, k_used + ko);
    o += (size_t)T * Np;
    ko += Np;
  }
  return 0;
}

int fedm_plan_resident_batches(const fedm_plan* P)
{
  int n = 0;
  for (const auto& g : P->groups)
    n += g.batch != nullptr && edm_batch_route(g.batch) == 1;
  return n;
}

int fedm_plan_device_batches(const fedm_plan* P)
{
  int n = 0;
  for (const auto& g : P->groups)
    n += g.batch != nullptr && edm_batch_on_device(g.batch);
  return n;
}

void fedm_plan_destroy(fedm_plan* P)
{
  delete P;
}

int fedm_run_command(edm_ctx* ctx, const fedm_command* c, fedm_result** out)
{
  *out = nullptr;
  fedm_plan* P = nullptr;
  int r = fedm_plan_create(ctx, c, &P);
  if (r == 0)
    r = fedm_plan_run(P);
  if (r == 0 || r == FEDM_BREAK_HIT) {
    const int q = fedm_plan_result(P, out);
    r = q != 0 ? q : r;
  }
  fedm_plan_destroy(P);
  return r;
}

int fedm_run_task_group(edm_ctx* ctx, const fedm_task_group* g, fedm_result** out)
{
  *out = nullptr;
  if (!ctx || !g || g->n <= 0 || g->num_E <= 0 || g->num_extras_lagged < 0 || g->num_extras_lagged > g->num_extras ||
      !g->t || !g->x || !g->obs || !g->usable || (g->num_extras > 0 && !g->extras))
    return -22;
  auto P = std::make_unique<fedm_plan>();
  P->ctx = ctx;
  const int n = g->n;
  Series& s = P->s;
  s.t.assign(g->t, g->t + n);
  s.x.assign(g->x, g->x + n);
  if (g->xmap) {
    s.xmap.assign(g->xmap, g->xmap + n);
    s.has_xmap = true;
  }
  if (g->co_x)
    s.co_x.assign(g->co_x, g->co_x + n);
  for (int e = 0; e < g->num_extras; e++)
    s.extras.emplace_back(g->extras + (size_t)e * n, g->extras + (size_t)(e + 1) * n);
  s.n_lagged = g->num_extras_lagged;
  s.obs.assign(g->obs, g->obs + n);
  if (g->panel) {
    s.panel.assign(g->panel, g->panel + n);
    s.has_panel = true;
  }
  s.tau = g->tau;
  s.p = g->p;
  s.dt = g->dt != 0;
  s.reldt = g->reldt != 0;
  s.w_dt = g->dt_weight;
  s.allow_missing = g->missing_distance != 0.0;

  edm_options& o = P->o;
  std::memset(&o, 0, sizeof(o));
  o.algorithm = g->algorithm;
  o.distance = g->distance;
  o.missing_distance = g->missing_distance;
  o.has_missing = s.allow_missing;
  o.force_compute = 1;
  o.save_smap_coeffs = g->save_smap_coeffs;
  o.aspect_ratio = g->aspect_ratio;
  o.precision = g->precision;
  P->thetas = g->num_thetas > 0 ? std::vector<double>(g->thetas, g->thetas + g->num_thetas) : std::vector<double>{ 1.0 };
  o.num_thetas = (int)P->thetas.size();
  o.thetas = P->thetas.data();
  o.idw = g->idw;
  o.panel_mode = g->panel_mode && g->panel != nullptr;
  P->save_smap_coeffs = g->save_smap_coeffs;
  if (g->extra_metrics) {
    P->extra_metrics.assign(g->extra_metrics, g->extra_metrics + g->num_extras);
    for (int m : P->extra_metrics)
      o.has_categorical |= m != 0;
  }

  std::vector<bool> usable(n);
  for (int i = 0; i < n; i++)
    usable[i] = g->usable[i] != 0;
  std::vector<int> libraries(g->libraries, g->libraries + g->num_libraries);
  const int nu = (int)std::count(usable.begin(), usable.end(), true);
  if (!g->explore && libraries.empty() && nu > 0)
    libraries = { nu };
  TaskGrid grid{ std::vector<int>(g->E, g->E + g->num_E), libraries, usable, g->k, g->num_reps, g->crossfold,
                 g->explore != 0, g->full != 0, g->shuffle != 0, g->save_final_predictions != 0,
                 g->save_final_copredictions != 0, g->copredict_mode != 0 && g->co_x != nullptr, g->seed };
  int r = plan_tasks(ctx, P.get(), grid);
  if (r == 0)
    r = fedm_plan_run(P.get());
  if (r == 0)
    r = fedm_plan_result(P.get(), out);
  return r;
}

int fedm_result_rc(const fedm_result* r)
{
  return r->rc;
}
int fedm_result_num_stats(const fedm_result* r, int co)
{
  return (int)(co ? r->costats.size() : r->stats.size());
}
void fedm_result_stats(const fedm_result* r, int co, int* E, int* library, double* theta, double* rho, double* mae)
{
  const auto& v = co ? r->costats : r->stats;
  for (size_t i = 0; i < v.size(); i++) {
    E[i] = v[i].E;
    library[i] = v[i].library;
    theta[i] = v[i].theta;
    rho[i] = v[i].rho;
    mae[i] = v[i].mae;
  }
}
void fedm_result_k(const fedm_result* r, int* kmin, int* kmax)
{
  *kmin = r->kmin;
  *kmax = r->kmax;
}
double fedm_result_scalar(const fedm_result* r, int which)
{
  return which == 0 ? r->missingdistance : r->dtWeight;
}
int fedm_result_matrix_shape(const fedm_result* r, int which, int* rows, int* cols)
{
  *rows = r->rows[which];
  *cols = r->cols[which];
  return r->has[which];
}
void fedm_result_matrix(const fedm_result* r, int which, double* out)
{
  std::copy(r->mat[which].begin(), r->mat[which].end(), out);
}
int fedm_result_num_manifolds(const fedm_result* r)
{
  return (int)r->Ms.size();
}
void fedm_result_manifold_shape(const fedm_result* r, int i, int ps, int* rows, int* cols)
{
  const Gathered& g = ps ? r->Mps[i] : r->Ms[i];
  *rows = g.n;
  *cols = g.Ea;
}
void fedm_result_manifold(const fedm_result* r, int i, int ps, double* out)
{
  const Gathered& g = ps ? r->Mps[i] : r->Ms[i];
  std::copy(g.data.begin(), g.data.end(), out);
}
void fedm_result_free(fedm_result* r)
{
  delete r;
}

} // extern "C"
