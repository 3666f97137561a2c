// Plain (FP64) kernels of the fastEDM prediction core for gfx950 (MI355X).
//
// Semantics follow the reference CPU path (EDM-Developers/fastEDM-r):
//   L^p distances        src/distances.cpp:151-229 (eager) and :69-149 (lazy/lowMemory)
//   k-NN order           src/edm.cpp:652-681  — ascending (distance, library position)
//   k policy             src/edm.cpp:593-621
//   simplex              src/edm.cpp:707-735
//   S-map                src/edm.cpp:737-814  — Gram X^T X, X^T y, then the JacobiSVD solve
// Built with -ffp-contract=off: every FP64 multiply and add is rounded separately, in the same
// order as the CPU path, so distances, neighbour order and S-map Gram matrices are bit-identical.
#include <algorithm>

#include "kernels_common.h"
#include "fdlibm_exp.h"

namespace fedm {

// ----------------------------------------------------------------------------------------------
// Pair distance of library point x and target y, coordinate j read by lib(j) and tgt(j).  Returns the
// accumulated distance, or MISSING_D when the pair is skipped or has a missing coordinate without
// missingdistance (both are dropped by the caller's keep rule).
template <int EMAX, bool GENERIC, class LIB, class TGT>
__device__ __forceinline__ double pair_acc_core(const LpParams& P, LIB lib, TGT tgt, int lib_panel, double lib_tt,
                                                int tgt_panel, double tgt_pt)
{
  double acc = 0.0;
  if (GENERIC) {
    const int f = P.flags;
    if ((f & F_SKIP_OTHER) && lib_panel != tgt_panel)
      return MISSING_D;
    if ((f & F_SKIP_FUTURE) && lib_tt > tgt_pt)
      return MISSING_D;
    if (f & F_IDW)
      acc += P.idw * (double)(lib_panel != tgt_panel);
    else if (f & F_IDWEIGHTS)
      acc += (double)P.id_weights[(size_t)lib_panel * P.num_panels + tgt_panel];
    const bool mae = f & F_MAE;
#pragma unroll
    for (int j = 0; j < EMAX; j++) {
      if (j < P.Ea) {
        const double x = lib(j), y = tgt(j);
        double d;
        if ((f & F_MISSING) && (x == MISSING_D || y == MISSING_D)) {
          if (P.missing_distance == 0.0)
            return MISSING_D;
          d = P.missing_distance;
        } else if ((f & F_CATEGORICAL) && P.metrics[j] == 1) {
          d = (double)(x != y);
        } else {
          d = x - y;
        }
        if (mae)
          acc += fabs(d) / (double)P.Ea;
        else
          acc += d * d;
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < EMAX; j++) {
      if (j < P.Ea) {
        const double d = lib(j) - tgt(j);
        acc += d * d;
      }
    }
  }
  return acc;
}

// The library point from LDS (xs), the target from registers (y).
template <int EMAX, bool GENERIC>
__device__ __forceinline__ double pair_acc(const LpParams& P, const double* __restrict__ xs, const double (&y)[EMAX],
                                           int i, int lib_panel, double lib_tt, int tgt_panel, double tgt_pt)
{
  return pair_acc_core<EMAX, GENERIC>(
    P, [&](int j) { return xs[j]; }, [&](int j) { return y[j]; }, lib_panel, lib_tt, tgt_panel, tgt_pt);
}

// Sorted insertion of (s, i) into a register-resident list of KCAP entries.
template <int KCAP>
__device__ __forceinline__ void list_insert(double (&ls)[KCAP], int (&li)[KCAP], double s, int i)
{
  double cs = s;
  int ci = i;
#pragma unroll
  for (int j = 0; j < KCAP; j++) {
    const bool take = key_less(cs, ci, ls[j], li[j]);
    const double ts = take ? cs : ls[j];
    const int ti = take ? ci : li[j];
    cs = take ? ls[j] : cs;
    ci = take ? li[j] : ci;
    ls[j] = ts;
    li[j] = ti;
  }
}

template <int KCAP>
__device__ __forceinline__ double list_at(const double (&ls)[KCAP], int k)
{
  double v = INF_D;
#pragma unroll
  for (int j = 0; j < KCAP; j++)
    if (j == k - 1)
      v = ls[j];
  return v;
}

// Staging of one library chunk into LDS (row-major [CH][Ea]), plus panel/target-time columns.
__device__ __forceinline__ void stage_chunk(const LpParams& P, double* lds_x, int* lds_panel, double* lds_tt, int c0,
                                            int cn, bool generic)
{
  const int total = cn * P.Ea;
  const double* src = P.M + (size_t)c0 * P.Ea;
  for (int e = threadIdx.x; e < total; e += blockDim.x)
    lds_x[e] = src[e];
  if (generic) {
    for (int e = threadIdx.x; e < cn; e += blockDim.x) {
      lds_panel[e] = P.Mpanel ? P.Mpanel[c0 + e] : 0;
      lds_tt[e] = P.Mtt ? P.Mtt[c0 + e] : 0.0;
    }
  }
}

// ----------------------------------------------------------------------------------------------
// Distance + exact top-KCAP per (target, library slice).
// grid = (ceil(NP / 256), S), block = 256: one thread per target, one library slice per blockIdx.y.
// Writes the slice's list sorted by (s, position) and its count of valid pairs.
template <int KCAP, int EMAX, bool GENERIC>
__global__ __launch_bounds__(256) void knn_slices_f64(LpParams P, int k, int slice_len, double* __restrict__ out_s,
                                                     int* __restrict__ out_i, int* __restrict__ out_cnt)
{
  __shared__ double lds_x[LDS_CHUNK_DOUBLES];
  __shared__ int lds_panel[LDS_CHUNK_MAXPTS];
  __shared__ double lds_tt[LDS_CHUNK_MAXPTS];

  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = t < P.NP;
  const int lo = blockIdx.y * slice_len;
  const int hi = min(P.NL, lo + slice_len);

  double y[EMAX];
#pragma unroll
  for (int j = 0; j < EMAX; j++)
    y[j] = (active && j < P.Ea) ? P.Mp[(size_t)t * P.Ea + j] : 0.0;
  const int tgt_panel = (GENERIC && active && P.Mppanel) ? P.Mppanel[t] : 0;
  const double tgt_pt = (GENERIC && active && P.Mppt) ? P.Mppt[t] : 0.0;

  double ls[KCAP];
  int li[KCAP];
#pragma unroll
  for (int j = 0; j < KCAP; j++) {
    ls[j] = INF_D;
    li[j] = INT_MAX;
  }
  double s_kth = INF_D, thr = INF_D;
  int cnt = 0;
  const bool mae = GENERIC && (P.flags & F_MAE);
  const int ch = min(LDS_CHUNK_MAXPTS, LDS_CHUNK_DOUBLES / P.Ea);

  for (int c0 = lo; c0 < hi; c0 += ch) {
    const int cn = min(ch, hi - c0);
    __syncthreads();
    stage_chunk(P, lds_x, lds_panel, lds_tt, c0, cn, GENERIC);
    __syncthreads();
    if (active) {
      for (int c = 0; c < cn; c++) {
        const double acc = pair_acc<EMAX, GENERIC>(P, lds_x + c * P.Ea, y, c0 + c, GENERIC ? lds_panel[c] : 0,
                                                   GENERIC ? lds_tt[c] : 0.0, tgt_panel, tgt_pt);
        if (acc != 0.0 && acc != MISSING_D) {
          cnt++;
          if (acc < thr) {
            const double s = mae ? acc : sqrt(acc);
            if (s < s_kth) {
              list_insert<KCAP>(ls, li, s, c0 + c);
              s_kth = list_at<KCAP>(ls, k);
              thr = (s_kth == INF_D) ? INF_D : (mae ? s_kth : s_kth * s_kth * SQRT_GUARD);
            }
          }
        }
      }
    }
  }
  if (active) {
    const size_t base = ((size_t)blockIdx.y * P.NP + t) * KCAP;
#pragma unroll
    for (int j = 0; j < KCAP; j++) {
      out_s[base + j] = ls[j];
      out_i[base + j] = li[j];
    }
    out_cnt[(size_t)blockIdx.y * P.NP + t] = cnt;
  }
}

// ----------------------------------------------------------------------------------------------
// Dense distances (k = all, or k wider than the register lists): D[t][i] = the distance s, or -1 where
// the pair is not a valid neighbour.  Library-major: lane = library point (its coordinates in
// registers), the block's DT_TGT targets staged in LDS, where every lane reads the same target at once.  A wavefront's stores of one target's row are 64 consecutive doubles.
// Block (x, y) takes library slice x (slice_len, a multiple of 256, points in tiles of 256) and targets
// y * DT_TGT .. + DT_TGT of the block of nt; out_cnt[x][t] = the slice's valid pairs of target t.
template <int EMAX, bool GENERIC>
__global__ __launch_bounds__(256) void dense_dist_f64(LpParams P, int t0, int nt, int slice_len,
                                                      double* __restrict__ D, int* __restrict__ out_cnt)
{
  __shared__ double ys[DT_TGT * EMAX];
  __shared__ int tpanel[DT_TGT];
  __shared__ double tpt[DT_TGT];
  __shared__ int wcnt[4][DT_TGT];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int tb = blockIdx.y * DT_TGT;
  const int ntb = min(DT_TGT, nt - tb);
  const int Ea = P.Ea;
  for (int e = tid; e < ntb * Ea; e += 256)
    ys[e] = P.Mp[(size_t)(t0 + tb) * Ea + e];
  if (tid < DT_TGT) {
    tpanel[tid] = (GENERIC && tid < ntb && P.Mppanel) ? P.Mppanel[t0 + tb + tid] : 0;
    tpt[tid] = (GENERIC && tid < ntb && P.Mppt) ? P.Mppt[t0 + tb + tid] : 0.0;
  }
  for (int e = tid; e < 4 * DT_TGT; e += 256)
    wcnt[e / DT_TGT][e % DT_TGT] = 0;
  __syncthreads();
  const bool mae = GENERIC && (P.flags & F_MAE);
  const int lo = blockIdx.x * slice_len, hi = min(P.NL, lo + slice_len);
  for (int i0 = lo; i0 < hi; i0 += 256) {
    const int i = i0 + tid;
    const bool inlib = i < hi;
    double x[EMAX];
#pragma unroll
    for (int j = 0; j < EMAX; j++)
      x[j] = (inlib && j < Ea) ? P.M[(size_t)i * Ea + j] : 0.0;
    const int lib_panel = (GENERIC && inlib && P.Mpanel) ? P.Mpanel[i] : 0;
    const double lib_tt = (GENERIC && inlib && P.Mtt) ? P.Mtt[i] : 0.0;
    for (int tl = 0; tl < ntb; tl++) {
      const double* yt = ys + tl * Ea;
      const double acc = pair_acc_core<EMAX, GENERIC>(
        P, [&](int j) { return x[j]; }, [&](int j) { return yt[j]; }, lib_panel, lib_tt, tpanel[tl], tpt[tl]);
      const bool v = inlib && acc != 0.0 && acc != MISSING_D;
      if (inlib)
        D[(size_t)(tb + tl) * P.NL + i] = v ? (mae ? acc : sqrt(acc)) : -1.0;
      const unsigned long long bb = __ballot(v);
      if (lane == 0)
        wcnt[wv][tl] += __popcll(bb);
    }
  }
  __syncthreads();
  if (tid < ntb)
    out_cnt[(size_t)blockIdx.x * P.NP + t0 + tb + tid] = wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
}

// ----------------------------------------------------------------------------------------------
// Dense distances for points wider than the register rows (E_actual > 64): pair_acc's arithmetic with
// the target's coordinates read from memory one at a time (they are shared by the block's threads only
// through the cache) and the library rows staged in LDS in chunks, one target per thread (grid: target
// blocks of 256 x library slices).  Same output as dense_dist_f64.
template <bool GENERIC>
__device__ __forceinline__ double pair_acc_rows(const LpParams& P, const double* __restrict__ xs,
                                                const double* __restrict__ y, int lib_panel, double lib_tt,
                                                int tgt_panel, double tgt_pt)
{
  double acc = 0.0;
  const int Ea = P.Ea;
  if (GENERIC) {
    const int f = P.flags;
    if ((f & F_SKIP_OTHER) && lib_panel != tgt_panel)
      return MISSING_D;
    if ((f & F_SKIP_FUTURE) && lib_tt > tgt_pt)
      return MISSING_D;
    if (f & F_IDW)
      acc += P.idw * (double)(lib_panel != tgt_panel);
    else if (f & F_IDWEIGHTS)
      acc += (double)P.id_weights[(size_t)lib_panel * P.num_panels + tgt_panel];
    const bool mae = f & F_MAE;
    for (int j = 0; j < Ea; j++) {
      const double x = xs[j], yj = y[j];
      double d;
      if ((f & F_MISSING) && (x == MISSING_D || yj == MISSING_D)) {
        if (P.missing_distance == 0.0)
          return MISSING_D;
        d = P.missing_distance;
      } else if ((f & F_CATEGORICAL) && P.metrics[j] == 1) {
        d = (double)(x != yj);
      } else {
        d = x - yj;
      }
      if (mae)
        acc += fabs(d) / (double)Ea;
      else
        acc += d * d;
    }
  } else {
    for (int j = 0; j < Ea; j++) {
      const double d = xs[j] - y[j];
      acc += d * d;
    }
  }
  return acc;
}

template <bool GENERIC>
__global__ __launch_bounds__(256) void dense_dist_rows(LpParams P, int t0, int nt, int slice_len,
                                                      double* __restrict__ D, int* __restrict__ out_cnt)
{
  __shared__ double lds_x[LDS_CHUNK_DOUBLES];
  __shared__ int lds_panel[LDS_CHUNK_MAXPTS];
  __shared__ double lds_tt[LDS_CHUNK_MAXPTS];

  const int tl = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = t0 + tl;
  const bool active = tl < nt && t < P.NP;
  const int lo = blockIdx.y * slice_len;
  const int hi = min(P.NL, lo + slice_len);
  const double* y = P.Mp + (size_t)(active ? t : 0) * P.Ea;
  const int tgt_panel = (GENERIC && active && P.Mppanel) ? P.Mppanel[t] : 0;
  const double tgt_pt = (GENERIC && active && P.Mppt) ? P.Mppt[t] : 0.0;
  const bool mae = GENERIC && (P.flags & F_MAE);
  const int ch = min(LDS_CHUNK_MAXPTS, LDS_CHUNK_DOUBLES / P.Ea);
  int cnt = 0;
  double* row = D + (size_t)tl * P.NL;
  for (int c0 = lo; c0 < hi; c0 += ch) {
    const int cn = min(ch, hi - c0);
    __syncthreads();
    stage_chunk(P, lds_x, lds_panel, lds_tt, c0, cn, GENERIC);
    __syncthreads();
    if (active) {
      for (int c = 0; c < cn; c++) {
        const double acc = pair_acc_rows<GENERIC>(P, lds_x + c * P.Ea, y, GENERIC ? lds_panel[c] : 0,
                                                  GENERIC ? lds_tt[c] : 0.0, tgt_panel, tgt_pt);
        double s = -1.0;
        if (acc != 0.0 && acc != MISSING_D) {
          cnt++;
          s = mae ? acc : sqrt(acc);
        }
        row[c0 + c] = s;
      }
    }
  }
  if (active)
    out_cnt[(size_t)blockIdx.y * P.NP + t] = cnt;
}

// ----------------------------------------------------------------------------------------------
// Merge the per-slice lists, apply the k policy (src/edm.cpp:593-621), record the neighbour set,
// and make the simplex predictions (src/edm.cpp:707-735).  One thread per target.
// nn_k[t] = number of neighbours in nn_i/nn_s; -1 means "all valid, in library order, see D";
// 0 means no prediction is made for this target.
template <int KCAP>
__global__ __launch_bounds__(256) void finish_lists(FinishParams F, const double* __restrict__ ps,
                                                   const int* __restrict__ pi, const int* __restrict__ pcnt, int S)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F.NP)
    return;
  int cnt = 0;
  for (int s = 0; s < S; s++)
    cnt += pcnt[(size_t)s * F.NP + t];

  int k = F.k;
  int* nk = F.nn_k + t;
  if (k > cnt) {
    if (F.force_compute) {
      k = cnt;
    } else {
      F.rc[t] = RC_INSUFFICIENT_UNIQUE;
      *nk = 0;
      return;
    }
  }
  if (k == 0 || cnt == 0) {
    F.rc[t] = F.force_compute ? RC_SUCCESS : RC_INSUFFICIENT_UNIQUE;
    *nk = 0;
    return;
  }
  int* oi = F.nn_i + (size_t)t * F.nn_stride;
  double* os = F.nn_s + (size_t)t * F.nn_stride;
  if (k == cnt) {
    // Every valid neighbour, in library order (src/edm.cpp:615-616). cnt <= KCAP here, so every
    // slice list holds all of its valid pairs; slices are in library order already.
    int o = 0;
    for (int s = 0; s < S; s++) {
      const int c = pcnt[(size_t)s * F.NP + t];
      // a slice list is sorted by distance: re-sort its c entries by position
      const size_t base = ((size_t)s * F.NP + t) * KCAP;
      for (int a = 0; a < c; a++) {
        oi[o + a] = pi[base + a];
        os[o + a] = ps[base + a];
      }
      for (int a = 1; a < c; a++) {
        const int ki = oi[o + a];
        const double ks = os[o + a];
        int b = a - 1;
        while (b >= 0 && oi[o + b] > ki) {
          oi[o + b + 1] = oi[o + b];
          os[o + b + 1] = os[o + b];
          b--;
        }
        oi[o + b + 1] = ki;
        os[o + b + 1] = ks;
      }
      o += c;
    }
  } else {
    // First k of the merged (s, position) order.
    int head[MAX_SLICES];
    for (int s = 0; s < S; s++)
      head[s] = 0;
    for (int o = 0; o < k; o++) {
      int best = -1;
      double bs = INF_D;
      int bi = INT_MAX;
      for (int s = 0; s < S; s++) {
        const int h = head[s];
        if (h >= KCAP)
          continue;
        const size_t base = ((size_t)s * F.NP + t) * KCAP + h;
        const double vs = ps[base];
        const int vi = pi[base];
        if (vi != INT_MAX && key_less(vs, vi, bs, bi)) {
          bs = vs;
          bi = vi;
          best = s;
        }
      }
      if (best < 0) {
        k = o;
        break;
      }
      head[best]++;
      oi[o] = bi;
      os[o] = bs;
    }
  }
  *nk = k;
  F.k_used[t] = k;
  if (F.nn_out) {
    for (int j = 0; j < F.nn_cap; j++)
      F.nn_out[(size_t)t * F.nn_cap + j] = j < k ? oi[j] : -1;
  }
  if (F.algorithm == ALG_SIMPLEX) {
    double minDist = os[0];
    for (int j = 1; j < k; j++)
      minDist = fmin(minDist, os[j]);
    for (int th = 0; th < F.T; th++) {
      const double theta = F.thetas[th];
      double sumw = 0.0;
      for (int j = 0; j < k; j++)
        sumw = sumw + fdlibm_exp(-theta * (os[j] / minDist));
      double r = 0.0;
      for (int j = 0; j < k; j++)
        r = r + F.targets[oi[j]] * (fdlibm_exp(-theta * (os[j] / minDist)) / sumw);
      F.pred[(size_t)th * F.NP + t] = r;
      F.rc[(size_t)th * F.NP + t] = RC_SUCCESS;
    }
  } else if (F.algorithm != ALG_SMAP) {
    F.rc[t] = RC_INVALID_ALGORITHM;
    *nk = 0;
  }
}

// Same policy for the dense route (k < 0, or k wider than the register lists): per-slice counts were
// written by dense_dist_f64.
__global__ __launch_bounds__(256) void finish_dense(FinishParams F, const double* __restrict__ D, int t0, int nt,
                                                   const int* __restrict__ pcnt, int S)
{
  const int tl = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = t0 + tl;
  if (tl >= nt || t >= F.NP)
    return;
  int cnt = 0;
  for (int s = 0; s < S; s++)
    cnt += pcnt[(size_t)s * F.NP + t];
  int k = F.k;
  if (k > cnt) {
    if (F.force_compute) {
      k = cnt;
    } else {
      F.rc[t] = RC_INSUFFICIENT_UNIQUE;
      F.nn_k[t] = 0;
      return;
    }
  }
  if (k == 0 || cnt == 0) {
    F.rc[t] = F.force_compute ? RC_SUCCESS : RC_INSUFFICIENT_UNIQUE;
    F.nn_k[t] = 0;
    return;
  }
  if (k > 0 && k < cnt) {
    // a selection wider than the register lists: select_wide makes it (and the simplex prediction);
    // the neighbour set is recorded for any algorithm, as the CPU path selects before it dispatches
    F.nn_cnt[t] = cnt;
    F.k_used[t] = k;
    F.nn_k[t] = k;
    if (F.algorithm != ALG_SIMPLEX && F.algorithm != ALG_SMAP)
      F.rc[t] = RC_INVALID_ALGORITHM;
    return;
  }
  // k < 0 or k == cnt: every valid neighbour in library order.
  F.nn_k[t] = -1;
  F.nn_cnt[t] = cnt;
  F.k_used[t] = cnt;
  const double* row = D + (size_t)tl * F.NL;
  if (F.nn_out) {
    int o = 0;
    for (int i = 0; i < F.NL && o < F.nn_cap; i++)
      if (row[i] >= 0.0)
        F.nn_out[(size_t)t * F.nn_cap + o++] = i;
    for (; o < F.nn_cap; o++)
      F.nn_out[(size_t)t * F.nn_cap + o] = -1;
  }
  if (F.algorithm == ALG_SIMPLEX) {
    double minDist = INF_D;
    for (int i = 0; i < F.NL; i++)
      if (row[i] >= 0.0)
        minDist = fmin(minDist, row[i]);
    for (int th = 0; th < F.T; th++) {
      const double theta = F.thetas[th];
      double sumw = 0.0;
      for (int i = 0; i < F.NL; i++)
        if (row[i] >= 0.0)
          sumw = sumw + fdlibm_exp(-theta * (row[i] / minDist));
      double r = 0.0;
      for (int i = 0; i < F.NL; i++)
        if (row[i] >= 0.0)
          r = r + F.targets[i] * (fdlibm_exp(-theta * (row[i] / minDist)) / sumw);
      F.pred[(size_t)th * F.NP + t] = r;
      F.rc[(size_t)th * F.NP + t] = RC_SUCCESS;
    }
  } else if (F.algorithm != ALG_SMAP) {
    F.rc[t] = RC_INVALID_ALGORITHM;
    F.nn_k[t] = 0;
  }
}

// ----------------------------------------------------------------------------------------------
// ----------------------------------------------------------------------------------------------
// Explicit instantiations used by the dispatcher.
#define FEDM_KNN(KC, EM)                                                                                         \
  template __global__ void knn_slices_f64<KC, EM, false>(LpParams, int, int, double*, int*, int*);               \
  template __global__ void knn_slices_f64<KC, EM, true>(LpParams, int, int, double*, int*, int*);
#define FEDM_KNN_E(KC) FEDM_KNN(KC, 8) FEDM_KNN(KC, 16) FEDM_KNN(KC, 32) FEDM_KNN(KC, 64)
FEDM_KNN_E(4)
FEDM_KNN_E(8)
FEDM_KNN_E(16)
FEDM_KNN_E(32)
FEDM_KNN_E(64)
template __global__ void dense_dist_f64<8, false>(LpParams, int, int, int, double*, int*);
template __global__ void dense_dist_f64<16, false>(LpParams, int, int, int, double*, int*);
template __global__ void dense_dist_f64<32, false>(LpParams, int, int, int, double*, int*);
template __global__ void dense_dist_f64<64, false>(LpParams, int, int, int, double*, int*);
template __global__ void dense_dist_f64<8, true>(LpParams, int, int, int, double*, int*);
template __global__ void dense_dist_f64<16, true>(LpParams, int, int, int, double*, int*);
template __global__ void dense_dist_f64<32, true>(LpParams, int, int, int, double*, int*);
template __global__ void dense_dist_f64<64, true>(LpParams, int, int, int, double*, int*);
template __global__ void dense_dist_rows<false>(LpParams, int, int, int, double*, int*);
template __global__ void dense_dist_rows<true>(LpParams, int, int, int, double*, int*);
template __global__ void finish_lists<4>(FinishParams, const double*, const int*, const int*, int);
template __global__ void finish_lists<8>(FinishParams, const double*, const int*, const int*, int);
template __global__ void finish_lists<16>(FinishParams, const double*, const int*, const int*, int);
template __global__ void finish_lists<32>(FinishParams, const double*, const int*, const int*, int);
template __global__ void finish_lists<64>(FinishParams, const double*, const int*, const int*, int);


} // namespace fedm
