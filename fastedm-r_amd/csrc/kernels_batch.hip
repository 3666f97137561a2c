// Batched simplex prediction over many tasks that share one point table (CCM library-size x
// replicate sweeps, E/theta grids): one persistent launch drains a work queue of (task, block of 256
// targets) items, heaviest libraries first.
//
// Per pair the library point is gathered through the task's row list into LDS, zero padded to EMAX
// coordinates (a zero coordinate adds exactly 0 to the sum, so the padded sum is the CPU path's sum,
// src/distances.cpp:186-216).  Two precisions:
//  * PLAIN  — FP64 distance;
//  * MIXED  — FP32 screening: the FP32 distance d32 is compared against the current k-th squared
//    distance plus a rigorous bound on |d32 - d64| (FP32 rounding of the inputs, of the differences
//    and of the FMA accumulation); only pairs that could enter the k-NN list are recomputed in FP64.
//
// The reference orders neighbours by (sqrt distance, library position) (src/edm.cpp:652-681).  The
// scan keeps the k smallest by (squared distance, position) instead — no square root per candidate,
// and since positions arrive in increasing order a new entry simply goes after equal squared
// distances.  The two orders differ only when two different squared distances have the same
// correctly rounded square root.  At the end the list is re-sorted by (sqrt, position); a pair left
// out of the k-th place whose squared distance lies strictly between the k-th's and the k-th's times
// (1 + 2^-49) could share the k-th's square root, so such a target is handed to batch_redo, which
// repeats it with the sqrt-ordered list of the CPU path.  The k policy (src/edm.cpp:593-621) and the
// simplex weights (src/edm.cpp:707-735) are the reference's.
#include "batch_common.h"
#include "fdlibm_exp.h"

namespace fedm {

// Sorted by (squared distance, position); the new entry has the largest position so far, so it goes
// after equal squared distances, and every entry behind it moves one place down.
template <int KCAP>
__device__ __forceinline__ void insert_acc(double (&la)[KCAP], int (&li)[KCAP], double a, int i)
{
  double cs = a;
  int ci = i;
  bool placed = false;
#pragma unroll
  for (int j = 0; j < KCAP; j++) {
    const bool take = placed || cs < la[j];
    placed = take;
    const double ts = take ? cs : la[j];
    const int ti = take ? ci : li[j];
    cs = take ? la[j] : cs;
    ci = take ? li[j] : ci;
    la[j] = ts;
    li[j] = ti;
  }
}

// Sorted by (sqrt distance, position).
template <int KCAP>
__device__ __forceinline__ void insert_sqrt(double (&ls)[KCAP], int (&li)[KCAP], double s, int i)
{
  double cs = s;
  int ci = i;
#pragma unroll
  for (int j = 0; j < KCAP; j++) {
    const bool take = cs < ls[j] || (cs == ls[j] && ci < li[j]);
    const double ts = take ? cs : ls[j];
    const int ti = take ? ci : li[j];
    cs = take ? ls[j] : cs;
    ci = take ? li[j] : ci;
    ls[j] = ts;
    li[j] = ti;
  }
}

template <int KCAP>
__device__ __forceinline__ double kth_of(const double (&ls)[KCAP], int k)
{
  double v = INF_D;
#pragma unroll
  for (int j = 0; j < KCAP; j++)
    if (j == k - 1)
      v = ls[j];
  return v;
}

// A float >= x, for x >= 0 (the next representable float when rounding went down).
__device__ __forceinline__ float float_up(double x)
{
  float f = (float)x;
  if ((double)f < x)
    f = __int_as_float(__float_as_int(f) + 1);
  return f;
}

// k policy, neighbour order and simplex prediction of target t of task tk.  ls/li: the neighbour list
// sorted by (sqrt distance, position); zeros: pairs dropped by the keep rule.
template <int KCAP>
__device__ void finish_target(const BatchParams& B, int tk, int t, int np, const int* lrows, int nl, int zeros,
                              int k, double (&ls)[KCAP], int (&li)[KCAP])
{
  const int cnt = nl - zeros;
  const int T = B.T;
  const int64_t ob = B.out_begin[tk];
  int ke = k;
  if (ke > cnt) {
    if (B.force_compute) {
      ke = cnt;
    } else {
      B.rc[ob + t] = RC_INSUFFICIENT_UNIQUE;
      return;
    }
  }
  if (ke == 0 || cnt == 0) {
    B.rc[ob + t] = B.force_compute ? RC_SUCCESS : RC_INSUFFICIENT_UNIQUE;
    return;
  }
  if (ke == cnt) {
    // every valid neighbour, in library order: sort the list by position
#pragma unroll
    for (int pass = 0; pass < KCAP; pass++) {
#pragma unroll
      for (int j = pass & 1; j + 1 < KCAP; j += 2) {
        const bool sw = li[j] > li[j + 1];
        const double a = ls[j], b = ls[j + 1];
        const int ia = li[j], ib = li[j + 1];
        ls[j] = sw ? b : a;
        ls[j + 1] = sw ? a : b;
        li[j] = sw ? ib : ia;
        li[j + 1] = sw ? ia : ib;
      }
    }
  }
  B.k_used[ob / T + t] = ke;
  double minDist = INF_D;
#pragma unroll
  for (int j = 0; j < KCAP; j++)
    if (j < ke)
      minDist = fmin(minDist, ls[j]);
  double tg[KCAP];
#pragma unroll
  for (int j = 0; j < KCAP; j++)
    tg[j] = j < ke ? B.Ltgt[lrows[li[j]]] : 0.0;
  for (int th = 0; th < T; th++) {
    const double theta = B.thetas[th];
    double w[KCAP];
    double sumw = 0.0;
#pragma unroll
    for (int j = 0; j < KCAP; j++)
      if (j < ke) {
        w[j] = fdlibm_exp(-theta * (ls[j] / minDist));
        sumw = sumw + w[j];
      }
    double r = 0.0;
#pragma unroll
    for (int j = 0; j < KCAP; j++)
      if (j < ke)
        r = r + tg[j] * (w[j] / sumw);
    B.pred[ob + (int64_t)th * np + t] = r;
    B.rc[ob + (int64_t)th * np + t] = RC_SUCCESS;
  }
}

constexpr int BCH = 512; // library points per LDS chunk

template <int KCAP, int EMAX, bool MIXED>
__global__ __launch_bounds__(256) void batch_simplex(BatchParams B)
{
  __shared__ double lds64[MIXED ? 1 : BCH * EMAX];
  __shared__ float lds32[MIXED ? BCH * EMAX : 1];
  __shared__ int lds_row[BCH];
  __shared__ int item_sh;

  const int Ea = B.Ea;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0)
      item_sh = atomicAdd(B.counter, 1);
    __syncthreads();
    const int item = item_sh;
    if (item >= B.num_items)
      break;
    const int2 it = B.items[item];
    const int tk = it.x;
    const int np = B.pred_count[tk];
    const int nl = B.lib_count[tk];
    const int t = it.y + threadIdx.x;
    const bool active = t < np;
    const int trow = active ? B.pred_rows[B.pred_begin[tk] + t] : 0;
    const int* lrows = B.lib_rows + B.lib_begin[tk];
    const int k = B.k[tk];

    double y[EMAX];
    float y32[MIXED ? EMAX : 1];
#pragma unroll
    for (int j = 0; j < EMAX; j++) {
      y[j] = (active && j < Ea) ? B.P[(size_t)trow * Ea + j] : 0.0;
      if (MIXED)
        y32[j] = (float)y[j];
    }
    double la[KCAP];
    int li[KCAP];
#pragma unroll
    for (int j = 0; j < KCAP; j++) {
      la[j] = INF_D;
      li[j] = INT_MAX;
    }
    double kth = INF_D, thr_g = INF_D;
    float thr32 = __builtin_huge_valf();
    int zeros = 0;
    bool near = false;

    for (int c0 = 0; c0 < nl; c0 += BCH) {
      const int cn = min(BCH, nl - c0);
      __syncthreads();
      for (int e = threadIdx.x; e < cn; e += blockDim.x)
        lds_row[e] = lrows[c0 + e];
      __syncthreads();
      for (int e = threadIdx.x; e < cn * EMAX; e += blockDim.x) {
        const int c = e / EMAX, j = e % EMAX;
        if (MIXED)
          lds32[e] = j < Ea ? B.L32[(size_t)lds_row[c] * Ea + j] : 0.0f;
        else
          lds64[e] = j < Ea ? B.L[(size_t)lds_row[c] * Ea + j] : 0.0;
      }
      __syncthreads();
      if (!active)
        continue;
      for (int c = 0; c < cn; c++) {
        double acc = 0.0;
        if (MIXED) {
          float d32 = 0.0f;
#pragma unroll
          for (int j = 0; j < EMAX; j++) {
            const float d = lds32[c * EMAX + j] - y32[j];
            d32 = __builtin_fmaf(d, d, d32);
          }
          if (d32 > thr32)
            continue;
          const double* xr = B.L + (size_t)lds_row[c] * Ea;
#pragma unroll
          for (int j = 0; j < EMAX; j++)
            if (j < Ea) {
              const double d = xr[j] - y[j];
              acc += d * d;
            }
        } else {
#pragma unroll
          for (int j = 0; j < EMAX; j++) {
            const double d = lds64[c * EMAX + j] - y[j];
            acc += d * d;
          }
        }
        if (acc == 0.0 || acc == MISSING_D) {
          zeros++;
          continue;
        }
        if (acc < thr_g) {
          if (acc < kth) {
            const double out = kth; // leaves the k-th place
            insert_acc<KCAP>(la, li, acc, c0 + c);
            kth = kth_of<KCAP>(la, k);
            if (out > kth && out < kth * SQRT_GUARD)
              near = true;
            if (kth != INF_D) {
              thr_g = kth * SQRT_GUARD;
              if (MIXED)
                thr32 = float_up(thr_g + 2.0 * (B.c1 * sqrt(thr_g) + B.c2 * thr_g + B.c3));
            }
          } else if (acc > kth) {
            near = true; // kth < acc < kth (1 + 2^-49): its square root may equal the k-th's
          }
        }
      }
    }
    if (!active)
      continue;
    if (near) {
      const int r = atomicAdd(B.redo_count, 1);
      if (r < B.redo_cap)
        B.redo[r] = make_int2(tk, t);
      continue;
    }
    double ls[KCAP];
    bool tie = false;
#pragma unroll
    for (int j = 0; j < KCAP; j++)
      ls[j] = la[j] == INF_D ? INF_D : sqrt(la[j]);
#pragma unroll
    for (int j = 0; j + 1 < KCAP; j++)
      tie = tie || (ls[j] == ls[j + 1] && li[j] > li[j + 1]);
    if (tie) {
      // equal square roots of different squared distances: order those by position
#pragma unroll
      for (int pass = 0; pass < KCAP; pass++) {
#pragma unroll
        for (int j = pass & 1; j + 1 < KCAP; j += 2) {
          const bool sw = ls[j + 1] < ls[j] || (ls[j + 1] == ls[j] && li[j + 1] < li[j]);
          const double a = ls[j], b = ls[j + 1];
          const int ia = li[j], ib = li[j + 1];
          ls[j] = sw ? b : a;
          ls[j + 1] = sw ? a : b;
          li[j] = sw ? ib : ia;
          li[j + 1] = sw ? ia : ib;
        }
      }
    }
    finish_target<KCAP>(B, tk, t, np, lrows, nl, zeros, k, ls, li);
  }
}

// The CPU path's selection for one (task, target): FP64 distances from the point table, list ordered
// by (sqrt distance, position) as they arrive.
template <int KCAP>
__device__ void exact_target(const BatchParams& B, int tk, int t)
{
  const int np = B.pred_count[tk];
  const int nl = B.lib_count[tk];
  const int Ea = B.Ea;
  const double* yp = B.P + (size_t)B.pred_rows[B.pred_begin[tk] + t] * Ea;
  const int* lrows = B.lib_rows + B.lib_begin[tk];
  const int k = B.k[tk];
  double ls[KCAP];
  int li[KCAP];
#pragma unroll
  for (int j = 0; j < KCAP; j++) {
    ls[j] = INF_D;
    li[j] = INT_MAX;
  }
  double s_kth = INF_D, thr = INF_D;
  int zeros = 0;
  // RB library rows at a time, so that their loads are in flight together; each pair's sum keeps
  // the coordinate order and the pairs are ranked in position order
  constexpr int RB = KCAP >= 32 ? 2 : KCAP >= 16 ? 4 : 8;
  for (int c0 = 0; c0 < nl; c0 += RB) {
    const double* xr[RB];
    double acc[RB];
#pragma unroll
    for (int u = 0; u < RB; u++) {
      xr[u] = B.L + (size_t)lrows[c0 + u < nl ? c0 + u : c0] * Ea;
      acc[u] = 0.0;
    }
    for (int j = 0; j < Ea; j++) {
      const double yj = yp[j];
#pragma unroll
      for (int u = 0; u < RB; u++) {
        const double d = xr[u][j] - yj;
        acc[u] += d * d;
      }
    }
#pragma unroll
    for (int u = 0; u < RB; u++) {
      const int c = c0 + u;
      if (c >= nl)
        break;
      if (acc[u] == 0.0 || acc[u] == MISSING_D) {
        zeros++;
        continue;
      }
      if (acc[u] < thr) {
        const double s = sqrt(acc[u]);
        if (s < s_kth) {
          insert_sqrt<KCAP>(ls, li, s, c);
          s_kth = kth_of<KCAP>(ls, k);
          if (s_kth != INF_D)
            thr = s_kth * s_kth * SQRT_GUARD;
        }
      }
    }
  }
  finish_target<KCAP>(B, tk, t, np, lrows, nl, zeros, k, ls, li);
}

// items == null: the (task, target) pairs handed over (at most redo_cap of them);
// otherwise every target of the given (task, first target) items (used when more were handed over).
template <int KCAP>
__global__ __launch_bounds__(256) void batch_redo(BatchParams B, const int2* items, int num_items)
{
  if (items) {
    for (int item = blockIdx.x; item < num_items; item += gridDim.x) {
      const int2 it = items[item];
      const int t = it.y + threadIdx.x;
      if (t < B.pred_count[it.x])
        exact_target<KCAP>(B, it.x, t);
    }
  } else {
    const int n = min(*B.redo_count, B.redo_cap);
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
      const int2 e = B.redo[r];
      exact_target<KCAP>(B, e.x, e.y);
    }
  }
}

#define FEDM_BATCH(KC, EM)                                                                                       \
  template __global__ void batch_simplex<KC, EM, false>(BatchParams);                                            \
  template __global__ void batch_simplex<KC, EM, true>(BatchParams);
#define FEDM_BATCH_E(KC)                                                                                         \
  FEDM_BATCH(KC, 4) FEDM_BATCH(KC, 8) FEDM_BATCH(KC, 16) FEDM_BATCH(KC, 32)                                      \
  template __global__ void batch_redo<KC>(BatchParams, const int2*, int);
FEDM_BATCH_E(4)
FEDM_BATCH_E(8)
FEDM_BATCH_E(16)
FEDM_BATCH_E(32)

} // namespace fedm
