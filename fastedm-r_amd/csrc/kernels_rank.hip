// Rank-table route for batches whose tasks share one prediction row list and draw their libraries
// from one row pool (CCM library-size x replicate sweeps).
//
// rank_build: for every target, the exact order of the pool rows by (sqrt(d64), row) — the reference's
//   k-NN order (src/edm.cpp:652-681), since library positions are monotone in rows — truncated to an
//   exact prefix of at most RANK_STORE rows, plus the rows at distance exactly 0 (dropped by the keep
//   rule, src/distances.cpp:218).  An FP32 histogram pass picks a cut; an FP32 filter with a rigorous
//   error bound collects every row whose FP64 distance can be below the cut; the FP64 distances of
//   those rows decide membership; the group tied with the largest kept distance is dropped, so the
//   kept rows are exactly {rows : s < s_max}; a bitonic sort in LDS orders them.
// rank_scan: a task's k nearest neighbours of a target are the first k rows of the target's ranked
//   prefix that belong to the task's library (a bitmask in LDS).  Targets whose prefix runs out, or
//   whose library is all-valid (k >= #valid), fall back to a direct exact scan of the library.
#include "batch_common.h"
#include "fdlibm_exp.h"

namespace fedm {

constexpr int RANK_BINS = 2048;   // FP32 bit-pattern bins: 8 exponent + 3 mantissa bits
constexpr int RANK_TARGET = 600;  // rows wanted below the histogram cut
constexpr int RANK_CAP = 1536;    // LDS candidate capacity
constexpr int RANK_SORT = 2048;   // bitonic sort width (power of two >= RANK_CAP)
__device__ __forceinline__ float up_float(double x)
{
  float f = (float)x;
  if ((double)f < x)
    f = __int_as_float(__float_as_int(f) + 1);
  return f;
}

template <int EMAX>
__global__ __launch_bounds__(256) void rank_build(RankParams R)
{
  __shared__ unsigned hist[RANK_BINS];
  __shared__ double cs[RANK_SORT];
  __shared__ int cq[RANK_SORT];
  __shared__ int ncand, nzero, overflow, cut_bin, all_rows;
  __shared__ double smax_sh;
  __shared__ int zr[RANK_ZCAP];

  const int t = blockIdx.x;
  if (t >= R.np)
    return;
  const int Ea = R.Ea;
  const int trow = R.pred[t];
  double y[EMAX];
  float y32[EMAX];
#pragma unroll
  for (int j = 0; j < EMAX; j++) {
    y[j] = j < Ea ? R.P[(size_t)trow * Ea + j] : 0.0;
    y32[j] = (float)y[j];
  }
  auto d32_of = [&](int row) {
    const float* x = R.L32 + (size_t)row * Ea;
    float d = 0.0f;
#pragma unroll
    for (int j = 0; j < EMAX; j++)
      if (j < Ea) {
        const float u = x[j] - y32[j];
        d = __builtin_fmaf(u, u, d);
      }
    return d;
  };
  auto d64_of = [&](int row) {
    const double* x = R.L + (size_t)row * Ea;
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < EMAX; j++)
      if (j < Ea) {
        const double u = x[j] - y[j];
        acc += u * u;
      }
    return acc;
  };

  // pass 1: histogram of the FP32 squared distances' bit patterns
  for (int b = threadIdx.x; b < RANK_BINS; b += blockDim.x)
    hist[b] = 0;
  __syncthreads();
  for (int q = threadIdx.x; q < R.npool; q += blockDim.x) {
    const float d = d32_of(R.pool[q]);
    atomicAdd(&hist[min(RANK_BINS - 1, (int)(__float_as_uint(d) >> 20))], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned run = 0;
    int b = 0;
    for (; b < RANK_BINS; b++) {
      run += hist[b];
      if (run >= (unsigned)RANK_TARGET)
        break;
    }
    all_rows = (R.npool <= RANK_CAP) ? 1 : 0;
    cut_bin = b;
  }
  __syncthreads();

  // pass 2: collect every row whose FP64 distance can be <= the cut
  for (int attempt = 0; attempt < 64; attempt++) {
    const int b = cut_bin;
    const bool all = all_rows != 0 || b >= RANK_BINS - 1;
    // cut tau: the lower edge of bin b+1 (every FP32 distance in bins <= b is below it)
    const double tau = all ? INF_D : (double)__uint_as_float((unsigned)(b + 1) << 20);
    const float T1 = all ? __builtin_huge_valf() : up_float(tau + 2.0 * (R.c1 * sqrt(tau) + R.c2 * tau + R.c3));
    if (threadIdx.x == 0) {
      ncand = 0;
      nzero = 0;
      overflow = 0;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < R.npool; q += blockDim.x) {
      const int row = R.pool[q];
      const float d = d32_of(row);
      if (!(d <= T1))
        continue;
      const double acc = d64_of(row);
      if (acc == 0.0) {
        const int z = atomicAdd(&nzero, 1);
        if (z < RANK_ZCAP)
          zr[z] = row;
      } else if (acc <= tau) {
        const int c = atomicAdd(&ncand, 1);
        if (c < RANK_CAP) {
          cs[c] = sqrt(acc);
          cq[c] = row;
        } else {
          overflow = 1;
        }
      }
    }
    __syncthreads();
    if (!overflow)
      break;
    __syncthreads();
    if (threadIdx.x == 0) {
      all_rows = 0;
      cut_bin = cut_bin - 1;
    }
    __syncthreads();
    if (cut_bin < 0)
      break;
  }
  if (overflow) {
    // no exact prefix fits: every task scans this target's library directly
    if (threadIdx.x == 0) {
      R.rank_len[t] = 0;
      R.zlen[t] = -1;
    }
    return;
  }
  const int n = min(ncand, RANK_CAP);
  // drop the group tied with the largest kept distance unless every pool row was examined
  if (!all_rows) {
    double m = 0.0;
    for (int c = threadIdx.x; c < n; c += blockDim.x)
      m = fmax(m, cs[c]);
    if (threadIdx.x == 0)
      smax_sh = 0.0;
    __syncthreads();
    // block max via LDS (values are non-negative doubles: compare as unsigned 64-bit)
    atomicMax((unsigned long long*)&smax_sh, (unsigned long long)__double_as_longlong(m));
    __syncthreads();
    const double smax = smax_sh;
    for (int c = threadIdx.x; c < n; c += blockDim.x)
      if (cs[c] == smax)
        cs[c] = INF_D;
  }
  int width = 64;
  while (width < n)
    width <<= 1;
  for (int c = n + threadIdx.x; c < width; c += blockDim.x) {
    cs[c] = INF_D;
    cq[c] = INT_MAX;
  }
  __syncthreads();
  // bitonic sort by (s, row) over the next power of two covering the candidates
  for (int size = 2; size <= width; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < width; i += blockDim.x) {
        const int j = i ^ stride;
        if (j > i) {
          const bool asc = (i & size) == 0;
          const double si = cs[i], sj = cs[j];
          const int qi = cq[i], qj = cq[j];
          const bool gt = si > sj || (si == sj && qi > qj);
          if (gt == asc) {
            cs[i] = sj;
            cs[j] = si;
            cq[i] = qj;
            cq[j] = qi;
          }
        }
      }
      __syncthreads();
    }
  }
  int kept = 0;
  if (threadIdx.x == 0) {
    int lo = 0, hi = n; // first INF position
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (cs[mid] != INF_D)
        lo = mid + 1;
      else
        hi = mid;
    }
    ncand = min(lo, RANK_STORE);
  }
  __syncthreads();
  kept = ncand;
  for (int c = threadIdx.x; c < kept; c += blockDim.x) {
    R.rank_row[(size_t)c * R.np + t] = cq[c];
    R.rank_s[(size_t)c * R.np + t] = cs[c];
  }
  if (threadIdx.x == 0) {
    R.rank_len[t] = kept;
    R.zlen[t] = nzero > RANK_ZCAP ? -1 : nzero;
  }
  if (threadIdx.x < RANK_ZCAP && threadIdx.x < nzero)
    R.zrow[(size_t)threadIdx.x * R.np + t] = zr[threadIdx.x];
}

template <int KCAP>
__device__ __forceinline__ void insert_sorted(double (&ls)[KCAP], int (&li)[KCAP], double s, int i)
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

template <int KCAP, int EMAX>
__global__ __launch_bounds__(256) void rank_scan(ScanParams S)
{
  extern __shared__ unsigned mask[];
  __shared__ int item_sh;
  const int Ea = S.Ea;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0)
      item_sh = atomicAdd(S.counter, 1);
    __syncthreads();
    const int item = item_sh;
    if (item >= S.num_items)
      break;
    const int2 it = S.items[item];
    const int tk = it.x;
    for (int w = threadIdx.x; w < S.mask_words; w += blockDim.x)
      mask[w] = S.masks[(size_t)tk * S.mask_words + w];
    __syncthreads();
    const int t = it.y + threadIdx.x;
    if (t >= S.np)
      continue;
    const int nl = S.lib_count[tk];
    const int k = S.k[tk];
    const int64_t ob = S.out_begin[tk];
    auto inlib = [&](int row) { return (mask[row >> 5] >> (row & 31)) & 1u; };

    // valid neighbours: every library row except those at distance 0
    const int zl = S.zlen[t];
    int zeros = 0;
    for (int z = 0; z < (zl < 0 ? 0 : zl); z++)
      zeros += inlib(S.zrow[(size_t)z * S.np + t]);
    const int cnt = nl - zeros;
    double ls[KCAP];
    int li[KCAP]; // table rows (ascending rows = ascending library positions)
#pragma unroll
    for (int j = 0; j < KCAP; j++) {
      ls[j] = INF_D;
      li[j] = INT_MAX;
    }
    int found = 0;
    bool direct = zl < 0 || k >= cnt;
    if (!direct) {
      const int len = S.rank_len[t];
      const int* rr = S.rank_row + t;
      const size_t st = (size_t)S.np;
      // four prefix entries per iteration keep several loads in flight
      int j = 0;
      for (; j + 4 <= len && found < k; j += 4) {
        const int r0 = rr[(size_t)j * st], r1 = rr[(size_t)(j + 1) * st], r2 = rr[(size_t)(j + 2) * st],
                  r3 = rr[(size_t)(j + 3) * st];
        const int rows4[4] = { r0, r1, r2, r3 };
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (found < k && inlib(rows4[u])) {
            const double s = S.rank_s[(size_t)(j + u) * st + t];
#pragma unroll
            for (int q = 0; q < KCAP; q++)
              if (q == found) {
                ls[q] = s;
                li[q] = rows4[u];
              }
            found++;
          }
        }
      }
      for (; j < len && found < k; j++) {
        const int row = rr[(size_t)j * st];
        if (inlib(row)) {
          const double s = S.rank_s[(size_t)j * st + t];
#pragma unroll
          for (int q = 0; q < KCAP; q++)
            if (q == found) {
              ls[q] = s;
              li[q] = row;
            }
          found++;
        }
      }
      direct = found < k;
    }
    int ke = k;
    if (direct) {
      // exact scan of the task's library (prefix exhausted, or k >= #valid)
#pragma unroll
      for (int j = 0; j < KCAP; j++) {
        ls[j] = INF_D;
        li[j] = INT_MAX;
      }
      const int trow = S.pred[t];
      double s_kth = INF_D, thr = INF_D;
      int zc = 0;
      const int* lr = S.lib_rows + S.lib_begin[tk];
      for (int c = 0; c < nl; c++) {
        const int row = lr[c];
        const double* x = S.L + (size_t)row * Ea;
        const double* yy = S.P + (size_t)trow * Ea;
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < EMAX; j++)
          if (j < Ea) {
            const double u = x[j] - yy[j];
            acc += u * u;
          }
        if (acc == 0.0 || acc == MISSING_D) {
          zc++;
          continue;
        }
        if (acc < thr) {
          const double s = sqrt(acc);
          if (s < s_kth) {
            insert_sorted<KCAP>(ls, li, s, row);
            double v = INF_D;
#pragma unroll
            for (int q = 0; q < KCAP; q++)
              if (q == k - 1)
                v = ls[q];
            s_kth = v;
            if (s_kth != INF_D)
              thr = s_kth * s_kth * SQRT_GUARD;
          }
        }
      }
      const int vcnt = nl - zc;
      if (ke > vcnt) {
        if (S.force_compute)
          ke = vcnt;
        else {
          S.rc[ob + t] = RC_INSUFFICIENT_UNIQUE;
          continue;
        }
      }
      if (ke == 0 || vcnt == 0) {
        S.rc[ob + t] = S.force_compute ? RC_SUCCESS : RC_INSUFFICIENT_UNIQUE;
        continue;
      }
      if (ke == vcnt) {
        // library order
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
    }
    S.k_used[ob / S.T + t] = ke;
    double minDist = INF_D;
    double tg[KCAP];
#pragma unroll
    for (int j = 0; j < KCAP; j++) {
      if (j < ke)
        minDist = fmin(minDist, ls[j]);
      tg[j] = j < ke ? S.Ltgt[li[j]] : 0.0;
    }
    for (int th = 0; th < S.T; th++) {
      const double theta = S.thetas[th];
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
      S.pred_out[ob + (int64_t)th * S.np + t] = r;
      S.rc[ob + (int64_t)th * S.np + t] = RC_SUCCESS;
    }
  }
}

#define FEDM_RANK_B(EM) template __global__ void rank_build<EM>(RankParams);
FEDM_RANK_B(4)
FEDM_RANK_B(8)
FEDM_RANK_B(16)
FEDM_RANK_B(32)
#define FEDM_RANK_S(KC, EM) template __global__ void rank_scan<KC, EM>(ScanParams);
#define FEDM_RANK_SE(KC) FEDM_RANK_S(KC, 4) FEDM_RANK_S(KC, 8) FEDM_RANK_S(KC, 16) FEDM_RANK_S(KC, 32)
FEDM_RANK_SE(4)
FEDM_RANK_SE(8)
FEDM_RANK_SE(16)
FEDM_RANK_SE(32)

} // namespace fedm
