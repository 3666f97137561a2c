// Batched simplex prediction over many tasks that share one point table (CCM library-size x
// replicate sweeps, E/theta grids): one persistent launch drains a work queue of (task, block of 256
// targets) items, heaviest libraries first.
//
// Per pair the library point is gathered through the task's row list into LDS.  Two precisions:
//  * PLAIN  — FP64 distance in the CPU path's operation order (src/distances.cpp:186-216);
//  * MIXED  — FP32 screening: the FP32 distance d32 is compared against the current k-th distance
//    plus a rigorous bound on |d32 - d64| (FP32 rounding of the inputs, of the differences and of the
//    FMA accumulation); only pairs that could enter the k-NN list are recomputed in FP64 and ranked
//    exactly by (sqrt(d64), library position).  The selected neighbours, and hence the predictions,
//    are identical to the FP64 route.
// The k-NN order (src/edm.cpp:652-681), the k policy (src/edm.cpp:593-621) and the simplex weights
// (src/edm.cpp:707-735) are those of the reference.
#include "kernels_common.h"
#include "fdlibm_exp.h"

namespace fedm {

struct BatchParams
{
  const double* L;      // library-side point table [NLrows x Ea] (FP64)
  const float* L32;     // FP32 copy of L (MIXED)
  const double* Ltgt;   // library-side targets [NLrows]
  const double* P;      // prediction-side point table [NProws x Ea]
  const int* lib_begin; // [tasks]
  const int* lib_count;
  const int* lib_rows;
  const int* pred_begin;
  const int* pred_count;
  const int* pred_rows;
  const int* k;          // [tasks]
  const int64_t* out_begin;
  const int2* items;     // (task, first target)
  int num_items;
  int* counter;          // work-queue head
  const double* thetas;
  int T, Ea, force_compute;
  double c1, c2, c3;     // mixed-precision error bound B(D) = c1 sqrt(D) + c2 D + c3
  double* pred;
  int* rc;
  int* k_used;
};

template <int KCAP>
__device__ __forceinline__ void list_insert_b(double (&ls)[KCAP], int (&li)[KCAP], double s, int i)
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
    double ls[KCAP];
    int li[KCAP];
#pragma unroll
    for (int j = 0; j < KCAP; j++) {
      ls[j] = INF_D;
      li[j] = INT_MAX;
    }
    double s_kth = INF_D, thr = INF_D;
    float thr32 = __builtin_huge_valf();
    int zeros = 0;

    for (int c0 = 0; c0 < nl; c0 += BCH) {
      const int cn = min(BCH, nl - c0);
      __syncthreads();
      for (int e = threadIdx.x; e < cn; e += blockDim.x)
        lds_row[e] = lrows[c0 + e];
      __syncthreads();
      for (int e = threadIdx.x; e < cn * Ea; e += blockDim.x) {
        const int c = e / Ea, j = e - c * Ea;
        if (MIXED)
          lds32[c * EMAX + j] = B.L32[(size_t)lds_row[c] * Ea + j];
        else
          lds64[c * EMAX + j] = B.L[(size_t)lds_row[c] * Ea + j];
      }
      __syncthreads();
      if (!active)
        continue;
      for (int c = 0; c < cn; c++) {
        bool exact = true;
        if (MIXED) {
          float d32 = 0.0f;
#pragma unroll
          for (int j = 0; j < EMAX; j++)
            if (j < Ea) {
              const float d = lds32[c * EMAX + j] - y32[j];
              d32 = __builtin_fmaf(d, d, d32);
            }
          exact = !(d32 > thr32);
        }
        if (!exact)
          continue;
        double acc = 0.0;
        if (MIXED) {
          const double* xr = B.L + (size_t)lds_row[c] * Ea;
#pragma unroll
          for (int j = 0; j < EMAX; j++)
            if (j < Ea) {
              const double d = xr[j] - y[j];
              acc += d * d;
            }
        } else {
#pragma unroll
          for (int j = 0; j < EMAX; j++)
            if (j < Ea) {
              const double d = lds64[c * EMAX + j] - y[j];
              acc += d * d;
            }
        }
        if (acc == 0.0 || acc == MISSING_D) {
          zeros++;
          continue;
        }
        if (acc < thr) {
          const double s = sqrt(acc);
          if (s < s_kth) {
            list_insert_b<KCAP>(ls, li, s, c0 + c);
            s_kth = kth_of<KCAP>(ls, k);
            if (s_kth != INF_D) {
              thr = s_kth * s_kth * SQRT_GUARD;
              if (MIXED) {
                const double bound = 2.0 * (B.c1 * sqrt(thr) + B.c2 * thr + B.c3);
                thr32 = float_up(thr + bound);
              }
            }
          }
        }
      }
    }
    if (!active)
      continue;

    // k policy (src/edm.cpp:593-621)
    const int cnt = nl - zeros;
    const int T = B.T;
    const int64_t ob = B.out_begin[tk];
    int ke = k;
    bool predict = true;
    if (ke > cnt) {
      if (B.force_compute)
        ke = cnt;
      else {
        B.rc[ob + t] = RC_INSUFFICIENT_UNIQUE;
        predict = false;
      }
    }
    if (predict && (ke == 0 || cnt == 0)) {
      B.rc[ob + t] = B.force_compute ? RC_SUCCESS : RC_INSUFFICIENT_UNIQUE;
      predict = false;
    }
    if (!predict)
      continue;
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
}

#define FEDM_BATCH(KC, EM)                                                                                       \
  template __global__ void batch_simplex<KC, EM, false>(BatchParams);                                            \
  template __global__ void batch_simplex<KC, EM, true>(BatchParams);
#define FEDM_BATCH_E(KC) FEDM_BATCH(KC, 4) FEDM_BATCH(KC, 8) FEDM_BATCH(KC, 16) FEDM_BATCH(KC, 32)
FEDM_BATCH_E(4)
FEDM_BATCH_E(8)
FEDM_BATCH_E(16)
FEDM_BATCH_E(32)

} // namespace fedm
