// Wasserstein ("curve matching") distances of the reference, src/distances.cpp:296-527, on the dense
// route: one wavefront per (target, library point) pair.
//
// The cost matrix of a pair (src/distances.cpp:296-439) is built in LDS, one entry per lane: the
// unlagged-extras and panel term fills every cell, then the |a - b| (Diff) or a != b (CheckSame)
// differences of each series row are added row by row, the dt row scaled by the plot aspect factor
// gamma.  Columns with a MISSING entry are dropped when missingdistance == 0 (so len_i != len_j is
// possible) or cost missingdistance otherwise.
//
// The reference solves the transport problem with uniform marginals 1/len_i, 1/len_j by network simplex
// (src/vendor/EMD_wrapper.h).  Its optimal value is unique, so any exact solver reproduces it up to the
// rounding of the final sum.  Here the wave runs successive shortest paths with node potentials on the
// bipartite graph (lanes are the len_i row nodes and len_j column nodes), in integer units of
// 1/(len_i len_j): each row supplies len_j units and each column takes len_i.  Every augmentation moves at
// least one unit, so the loop ends after at most len_i len_j augmentations.  The cost is then summed as
// flow * C in the reference's arc order (last arc id first, src/vendor/EMD_wrapper.h:106-113).  A pair is
// kept iff its distance is nonzero and normal (src/distances.cpp:515).
#include "kernels_common.h"

namespace fedm {

// NPL transport nodes per lane: len_i + len_j <= 64 NPL, so series of up to 32 NPL observations.  NPL = 1
// takes E_x <= 32 with four wavefronts per block; NPL = 2 takes E_x <= 64 with two (48 KB of LDS each).

__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// (value, lane) minimum over the wave; ties go to the lower lane.
__device__ __forceinline__ void wave_argmin(double& v, int& idx)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(v, o);
    const int oi = __shfl_xor(idx, o);
    if (ov < v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
}

// (wave_max comes with the shared kernel header)
__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v = fmin(v, __shfl_xor(v, o));
  return v;
}

// Exact optimal transport value of the li x lj cost C (row-major, LDS) with uniform marginals.
// F: li x lj integer flows (LDS, zeroed by the caller).  Node u (rows 0 .. li - 1, then columns) lives on
// lane u % 64, slot u / 64.  Returns NaN if no augmenting path is found (impossible for finite costs), so
// the pair is not kept.
template <int NPL>
__device__ double transport_wave(const double* C, int* F, int li, int lj, int lane)
{
  const int nn = li + lj;
  bool isrow[NPL], iscol[NPL];
  int jc[NPL], sup[NPL];
  double pot[NPL];
#pragma unroll
  for (int s = 0; s < NPL; s++) {
    const int u = lane + 64 * s;
    isrow[s] = u < li;
    iscol[s] = u >= li && u < nn;
    jc[s] = u - li;
    pot[s] = 0.0;
    if (iscol[s]) {
      double mn = INF_D;
      for (int i = 0; i < li; i++)
        mn = fmin(mn, C[i * lj + jc[s]]);
      pot[s] = mn < 0.0 ? -mn : 0.0; // reduced costs C - pot_row + pot_col start non-negative
    }
    sup[s] = isrow[s] ? lj : (iscol[s] ? li : 0);
  }
  // a node's value from its owner lane (u wave-uniform)
  auto at = [&](const auto (&v)[NPL], int u) {
    auto x = v[0];
#pragma unroll
    for (int s = 1; s < NPL; s++)
      if (u >> 6 == s)
        x = v[s];
    return __shfl(x, u & 63);
  };
  int remaining = li * lj;
  while (remaining > 0) {
    // the first row node with supply left
    int src = -1;
#pragma unroll
    for (int s = NPL - 1; s >= 0; s--) {
      const unsigned long long sm = __ballot(isrow[s] && sup[s] > 0);
      if (sm)
        src = 64 * s + __ffsll((long long)sm) - 1;
    }
    double d[NPL];
    bool done[NPL];
    int pred[NPL];
#pragma unroll
    for (int s = 0; s < NPL; s++) {
      d[s] = lane + 64 * s == src ? 0.0 : INF_D;
      done[s] = !(isrow[s] || iscol[s]);
      pred[s] = -1;
    }
    int tgt = -1;
    for (int it = 0; it < nn; it++) {
      // (value, node) minimum over the not yet settled nodes; ties go to the lower node
      double v = INF_D;
      int u = 0x7fffffff;
#pragma unroll
      for (int s = 0; s < NPL; s++) {
        const double vs = done[s] ? INF_D : d[s];
        if (vs < v) {
          v = vs;
          u = lane + 64 * s;
        }
      }
      if (v == INF_D)
        u = lane; // an empty lane takes its own id, as in the one-node case
      wave_argmin(v, u);
      if (v == INF_D)
        break;
#pragma unroll
      for (int s = 0; s < NPL; s++)
        if (lane + 64 * s == u)
          done[s] = true;
      if (u >= li && at(sup, u) > 0) {
        tgt = u;
        break;
      }
      const double pu = at(pot, u);
      if (u < li) {
#pragma unroll
        for (int s = 0; s < NPL; s++)
          if (iscol[s] && !done[s]) {
            const double cand = v + (C[u * lj + jc[s]] - pu + pot[s]);
            if (cand < d[s]) {
              d[s] = cand;
              pred[s] = u;
            }
          }
      } else {
        const int ju = u - li;
#pragma unroll
        for (int s = 0; s < NPL; s++) {
          const int r = lane + 64 * s;
          if (isrow[s] && !done[s] && F[r * lj + ju] > 0) {
            const double cand = v + (-C[r * lj + ju] - pu + pot[s]);
            if (cand < d[s]) {
              d[s] = cand;
              pred[s] = u;
            }
          }
        }
      }
    }
    if (tgt < 0)
      return __builtin_nan("");
    const double dt = at(d, tgt);
#pragma unroll
    for (int s = 0; s < NPL; s++)
      if (isrow[s] || iscol[s])
        pot[s] -= fmin(d[s], dt);
    // amount: remaining supply at src, remaining demand at tgt, flow on the backward arcs of the path
    int delta = min(at(sup, src), at(sup, tgt));
    int v = tgt;
    for (int s = 0; s < nn && v != src; s++) {
      const int p = at(pred, v);
      if (p < 0)
        return __builtin_nan("");
      if (p >= li)
        delta = min(delta, F[v * lj + (p - li)]);
      v = p;
    }
    if (v != src || delta <= 0)
      return __builtin_nan("");
    v = tgt;
    for (int s = 0; s < nn && v != src; s++) {
      const int p = at(pred, v);
      if (lane == 0) {
        if (p < li)
          F[p * lj + (v - li)] += delta;
        else
          F[v * lj + (p - li)] -= delta;
      }
      v = p;
    }
    wave_sync();
#pragma unroll
    for (int s = 0; s < NPL; s++) {
      if (lane + 64 * s == src)
        sup[s] -= delta;
      if (lane + 64 * s == tgt)
        sup[s] -= delta;
    }
    remaining -= delta;
  }
  double cost = 0.0;
  if (lane == 0) {
    const double units = (double)li * (double)lj;
    for (int a = li * lj - 1; a >= 0; a--) {
      const int f = F[a];
      if (f > 0)
        cost += ((double)f / units) * C[a];
    }
  }
  return __shfl(cost, 0);
}

// Block (target t0 + blockIdx.x, library slice blockIdx.y); wave w takes library points lo + w, lo + w + 4, ...
// Writes the distance row (kept distances, -1 otherwise) and the slice's kept count.
template <int NPL, int W_WAVES>
__global__ __launch_bounds__(64 * W_WAVES) void wasserstein_dense(WassParams P, int t0, int nt, int slice_len,
                                                                 double* __restrict__ D, int* __restrict__ out_cnt)
{
  constexpr int W_EMAX = 32 * NPL; // longest series: len_i + len_j <= 64 NPL nodes
  __shared__ double Cs[W_WAVES][W_EMAX * W_EMAX];
  __shared__ int Fs[W_WAVES][W_EMAX * W_EMAX];
  __shared__ int cmi[W_WAVES][W_EMAX];
  __shared__ int cmj[W_WAVES][W_EMAX];
  __shared__ int kept_s;

  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tl = blockIdx.x, t = t0 + tl;
  if (tl >= nt || t >= P.NP)
    return;
  const int lo = blockIdx.y * slice_len;
  const int hi = min(P.NL, lo + slice_len);
  if (threadIdx.x == 0)
    kept_s = 0;
  __syncthreads();

  const int E = P.E, rows = P.rows;
  const unsigned long long emask = E >= 64 ? ~0ull : ((1ull << E) - 1ull);
  const unsigned long long below = (1ull << lane) - 1ull;
  const bool skip_missing = P.missing_distance == 0.0;
  const double* y = P.Mp + (size_t)t * P.Ea;
  const int tp = P.Mppanel ? P.Mppanel[t] : 0;

  bool mj = false;
  if (lane < E)
    for (int r = 0; r < rows; r++)
      mj = mj || y[r * E + lane] == MISSING_D;
  const unsigned long long mj_mask = __ballot(mj) & emask;
  const unsigned long long keep_j = skip_missing ? (~mj_mask & emask) : emask;
  const int lj = __popcll(keep_j);
  if (lane < E && ((keep_j >> lane) & 1ull))
    cmj[w][__popcll(keep_j & below)] = lane;

  int kept = 0;
  double* row = D + (size_t)tl * P.NL;
  for (int i = lo + w; i < hi; i += W_WAVES) {
    const double* x = P.M + (size_t)i * P.Ea;
    double dist = -1.0;
    if (!(P.skip_other && P.Mpanel[i] != tp)) {
      bool mi = false;
      if (lane < E)
        for (int r = 0; r < rows; r++)
          mi = mi || x[r * E + lane] == MISSING_D;
      const unsigned long long mi_mask = __ballot(mi) & emask;
      const unsigned long long keep_i = skip_missing ? (~mi_mask & emask) : emask;
      const int li = __popcll(keep_i);
      if (li > 0 && lj > 0) {
        double gamma = 1.0;
        if (P.E_dt > 0) {
          // plot aspect ratio of the library curve (src/distances.cpp:344-366)
          double mn = DBL_MAX, mx = DBL_MIN, mt = 0.0;
          if (lane < E) {
            const double v = x[lane];
            if (v != MISSING_D) {
              mn = fmin(mn, v);
              mx = fmax(mx, v);
            }
            const double tt = x[E + lane];
            if (tt != MISSING_D && tt > mt)
              mt = tt;
          }
          mn = wave_min(mn);
          mx = wave_max(mx);
          mt = wave_max(mt);
          gamma = P.aspect_ratio * (mx - mn + 1e-6) / (mt + 1e-6);
        }
        double unl = 0.0;
        for (int e = 0; e < P.n_unlagged; e++) {
          const double xe = x[rows * E + e], ye = y[rows * E + e];
          if (xe == MISSING_D || ye == MISSING_D)
            unl += P.missing_distance;
          else if (P.metrics[rows + e] == 0)
            unl += fabs(xe - ye);
          else
            unl += (xe != ye) ? 1.0 : 0.0;
        }
        if (P.panel_term == 1)
          unl += P.idw * ((P.Mpanel[i] != tp) ? 1.0 : 0.0);
        else if (P.panel_term == 2)
          unl += (double)P.id_weights[(size_t)P.Mpanel[i] * P.num_panels + tp];
        if (lane < E && ((keep_i >> lane) & 1ull))
          cmi[w][__popcll(keep_i & below)] = lane;
        wave_sync();
        for (int e = lane; e < li * lj; e += 64) {
          const int n = e / lj, m = e - n * lj;
          const int a = cmi[w][n], b = cmj[w][m];
          const bool miss = ((mi_mask >> a) & 1ull) || ((mj_mask >> b) & 1ull);
          double c = unl;
          for (int k = 0; k < rows; k++) {
            double dk;
            if (miss)
              dk = P.missing_distance;
            else if (P.metrics[k] == 0)
              dk = fabs(x[k * E + a] - y[k * E + b]);
            else
              dk = (x[k * E + a] != y[k * E + b]) ? 1.0 : 0.0;
            if (P.E_dt > 0 && k == 1)
              dk *= gamma;
            c += dk;
          }
          Cs[w][e] = c;
          Fs[w][e] = 0;
        }
        wave_sync();
        const double cost = transport_wave<NPL>(Cs[w], Fs[w], li, lj, lane);
        if (cost != 0.0 && fabs(cost) >= DBL_MIN && fabs(cost) < INF_D)
          dist = cost;
        wave_sync();
      }
    }
    if (lane == 0) {
      row[i] = dist;
      kept += dist >= 0.0 ? 1 : 0;
    }
  }
  if (lane == 0)
    atomicAdd(&kept_s, kept);
  __syncthreads();
  if (threadIdx.x == 0)
    out_cnt[(size_t)blockIdx.y * P.NP + t] = kept_s;
}

template __global__ void wasserstein_dense<1, 4>(WassParams, int, int, int, double*, int*);
template __global__ void wasserstein_dense<2, 2>(WassParams, int, int, int, double*, int*);

} // namespace fedm
