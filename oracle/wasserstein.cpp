/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY.  Wasserstein ("curve matching") distance of the reference,
 * src/distances.cpp:296-527, with the exact optimal-transport value that src/vendor/EMD_wrapper.h
 * (network simplex, uniform marginals 1/len_i and 1/len_j) computes.  The transport problem is solved
 * here exactly by successive shortest paths on integer-scaled supplies (each source supplies len_j
 * units, each sink demands len_i units); the optimal cost is unique, so any exact solver gives the
 * reference's value up to the rounding of the final sum.  The reference's unqualified abs() on doubles
 * is taken as |.| (parity of that quirk is unpinned).
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace orc {

const double W_MISSING_D = 1.0e+100;

struct Pairs
{
  std::vector<int> inds;
  std::vector<double> dists;
};

// Exact min-cost transport with uniform marginals (integer-scaled) by successive shortest paths.
double exact_emd(const double* C, int li, int lj)
{
  const int S = li + lj + 2, src = li + lj, snk = li + lj + 1;
  struct Edge
  {
    int to, rev;
    long cap;
    double cost;
  };
  std::vector<std::vector<Edge>> g(S);
  auto add = [&](int a, int b, long cap, double cost) {
    g[a].push_back({ b, (int)g[b].size(), cap, cost });
    g[b].push_back({ a, (int)g[a].size() - 1, 0, -cost });
  };
  for (int i = 0; i < li; i++)
    add(src, i, lj, 0.0);
  for (int j = 0; j < lj; j++)
    add(li + j, snk, li, 0.0);
  for (int i = 0; i < li; i++)
    for (int j = 0; j < lj; j++)
      add(i, li + j, (long)li * lj, C[(size_t)i * lj + j]);
  long need = (long)li * lj;
  double total = 0.0;
  while (need > 0) {
    std::vector<double> dist(S, std::numeric_limits<double>::infinity());
    std::vector<int> pv(S, -1), pe(S, -1);
    dist[src] = 0;
    for (int it = 0; it < S; it++) {
      bool upd = false;
      for (int u = 0; u < S; u++) {
        if (dist[u] == std::numeric_limits<double>::infinity())
          continue;
        for (int e = 0; e < (int)g[u].size(); e++) {
          const Edge& ed = g[u][e];
          const double cand = dist[u] + ed.cost;
          const double cur = dist[ed.to];
          if (ed.cap > 0 && (std::isinf(cur) || cand < cur - 1e-15 * (1.0 + std::fabs(cur)))) {
            dist[ed.to] = cand;
            pv[ed.to] = u;
            pe[ed.to] = e;
            upd = true;
          }
        }
      }
      if (!upd)
        break;
    }
    if (pv[snk] < 0)
      break;
    long f = need;
    for (int v = snk; v != src; v = pv[v])
      f = std::min(f, g[pv[v]][pe[v]].cap);
    for (int v = snk; v != src; v = pv[v]) {
      Edge& ed = g[pv[v]][pe[v]];
      ed.cap -= f;
      g[v][ed.rev].cap += f;
    }
    need -= f;
  }
  // Flow on each (i,j) arc in units of 1/(li*lj), summed as flow_ij * C_ij in the reference's arc order:
  // EMD_wrap walks the arcs from the last id (i*lj + j) down to 0 (src/vendor/EMD_wrapper.h:106-113,
  // src/vendor/full_bipartitegraph.h:98-100).
  std::vector<long> flow((size_t)li * lj, 0);
  for (int i = 0; i < li; i++)
    for (const Edge& ed : g[i])
      if (ed.to >= li && ed.to < li + lj)
        flow[(size_t)i * lj + (ed.to - li)] = (long)li * lj - ed.cap;
  for (long a = (long)li * lj - 1; a >= 0; a--)
    if (flow[a] > 0)
      total += ((double)flow[a] / ((double)li * lj)) * C[a];
  return total;
}

// src/distances.cpp:296-439
static std::vector<double> cost_matrix(const orc_manifold_view& M, const orc_manifold_view& Mp, int i, int j,
                                       const orc_options& o, int& len_i, int& len_j)
{
  const bool skipMissing = (o.missing_distance == 0);
  const int E = M.E_x, Ea = M.E_actual;
  const double* x = M.data + (size_t)i * Ea;
  const double* y = Mp.data + (size_t)j * Ea;
  const int nLaggedExtras = M.E_lagged_extras / E;
  const int rows = 1 + (M.E_dt > 0) + nLaggedExtras;
  std::vector<char> mi(E, 0), mj(E, 0);
  for (int c = 0; c < E; c++)
    for (int r = 0; r < rows; r++) {
      if (x[r * E + c] == W_MISSING_D)
        mi[c] = 1;
      if (y[r * E + c] == W_MISSING_D)
        mj[c] = 1;
    }
  if (skipMissing) {
    len_i = E - (int)std::count(mi.begin(), mi.end(), 1);
    len_j = E - (int)std::count(mj.begin(), mj.end(), 1);
  } else {
    len_i = E;
    len_j = E;
  }
  double gamma = 1.0;
  if (M.E_dt > 0) {
    double minData = std::numeric_limits<double>::max();
    double maxData = std::numeric_limits<double>::min();
    double maxTime = 0.0;
    for (int c = 0; c < E; c++) {
      double v = x[c];
      if (v != W_MISSING_D) {
        if (v < minData)
          minData = v;
        if (v > maxData)
          maxData = v;
      }
      double tt = x[E + c];
      if (tt != W_MISSING_D && tt > maxTime)
        maxTime = tt;
    }
    double eps = 1e-6;
    gamma = o.aspect_ratio * (maxData - minData + eps) / (maxTime + eps);
  }
  double unlagged = 0.0;
  int nUnlagged = M.E_extras - M.E_lagged_extras;
  for (int e = 0; e < nUnlagged; e++) {
    double xe = x[rows * E + e], ye = y[rows * E + e];
    if (xe == W_MISSING_D || ye == W_MISSING_D)
      unlagged += o.missing_distance;
    else if (o.metrics == nullptr || o.metrics[rows + e] == 0)
      unlagged += std::fabs(xe - ye);
    else
      unlagged += (xe != ye);
  }
  if (o.panel_mode && o.idw > 0)
    unlagged += o.idw * (M.panel[i] != Mp.panel[j]);
  else if (o.panel_mode && o.id_weights != nullptr)
    unlagged += (double)o.id_weights[(size_t)M.panel[i] * o.num_panels + Mp.panel[j]];

  std::vector<double> C((size_t)std::max(0, len_i * len_j), unlagged);
  for (int k = 0; k < rows; k++) {
    int n = 0;
    for (int a = 0; a < E; a++) {
      if (skipMissing && mi[a])
        continue;
      int m = 0;
      for (int b = 0; b < E; b++) {
        if (skipMissing && mj[b])
          continue;
        double d;
        if (mi[a] || mj[b])
          d = o.missing_distance;
        else if (o.metrics == nullptr || o.metrics[k] == 0)
          d = std::fabs(x[k * E + a] - y[k * E + b]);
        else
          d = x[k * E + a] != y[k * E + b];
        if (M.E_dt > 0 && k == 1)
          d *= gamma;
        C[(size_t)n * len_j + m] += d;
        m++;
      }
      n++;
    }
  }
  return C;
}

// src/distances.cpp:490-527
Pairs wasserstein_distances(int Mp_i, const orc_options& o, const orc_manifold_view& M, const orc_manifold_view& Mp)
{
  Pairs out;
  const bool skipOther = o.panel_mode && o.idw < 0;
  for (int i = 0; i < M.num_points; i++) {
    if (skipOther && M.panel[i] != Mp.panel[Mp_i])
      continue;
    int li, lj;
    std::vector<double> C = cost_matrix(M, Mp, i, Mp_i, o, li, lj);
    if (li > 0 && lj > 0) {
      double d = exact_emd(C.data(), li, lj);
      if (d != 0 && std::isnormal(d)) {
        out.dists.push_back(d);
        out.inds.push_back(i);
      }
    }
  }
  return out;
}

} // namespace orc

extern "C" double orc_emd(const double* C, int li, int lj)
{
  return orc::exact_emd(C, li, lj);
}
