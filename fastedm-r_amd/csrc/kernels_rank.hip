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
constexpr int RANK_TARGET = 500;  // rows wanted below the histogram cut
constexpr int RANK_SAMPLE = 8;    // the histogram counts one pool row in RANK_SAMPLE (the hash picks 0..7)
constexpr int RANK_SAMPLE_TARGET = (RANK_TARGET + RANK_SAMPLE - 1) / RANK_SAMPLE;
constexpr int RANK_CAP = 1024;    // LDS candidate capacity (the stored prefix is at most RANK_STORE)
constexpr int RANK_SORT = 1024;   // bitonic sort width (power of two >= RANK_CAP)
constexpr int RANK_NB = 256;      // buckets of the counting sort of the candidates (by s in [0, s_max])
constexpr int RANK_BMAX = 32;     // a bucket holding more candidates sends the target to the bitonic sort
static_assert(RANK_STORE <= RANK_CAP, "the stored prefix is cut from the sorted candidates");
static_assert(RANK_BINS * sizeof(unsigned) <= RANK_SORT * sizeof(double), "the histogram shares the sort keys' LDS");
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
  // The histogram is done with before pass 2 writes the first candidate, so it shares the sort keys'
  // storage; with the counting sort's buckets and order, about 18 KiB of LDS per block, eight blocks per
  // CU.
  __shared__ double cs[RANK_SORT];
  unsigned* hist = reinterpret_cast<unsigned*>(cs);
  __shared__ int cq[RANK_SORT];
  __shared__ int ord[RANK_CAP];
  __shared__ unsigned bcnt[RANK_NB], bstart[RANK_NB + 1];
  __shared__ unsigned bmaxc;
  __shared__ int ncand, nzero, overflow, cut_bin, all_rows, nlist;
  __shared__ unsigned wsum[4];
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
  // FP32 squared distance of pool entry q (padded coordinates are 0 on both sides and add 0)
  auto d32_of = [&](int q) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f* x = (const v4f*)(R.pool32 + (size_t)q * EMAX);
    float d = 0.0f;
#pragma unroll
    for (int j4 = 0; j4 < EMAX / 4; j4++) {
      const v4f v = x[j4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const float w = v[u] - y32[4 * j4 + u];
        d = __builtin_fmaf(w, w, d);
      }
    }
    return d;
  };
  constexpr int UNR = 4; // pool entries per thread whose loads are in flight together
  auto d64_of = [&](int row) {
    // every coordinate load is issued at a clamped index before the sum, so that they are in flight
    // together; a padding coordinate adds +0.0, which leaves the sum (and its order) unchanged
    const double* x = R.L + (size_t)row * Ea;
    double v[EMAX];
#pragma unroll
    for (int j = 0; j < EMAX; j++)
      v[j] = x[j < Ea ? j : Ea - 1];
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < EMAX; j++) {
      const double u = j < Ea ? v[j] - y[j] : 0.0;
      acc += u * u;
    }
    return acc;
  };

  // pass 1: histogram of the FP32 squared distances' bit patterns over one pool row in each run of
  // RANK_SAMPLE, at a hashed offset (a fixed stride would alias with periodic series).  The cut only
  // sizes the prefix (any cut gives an exact prefix): the sample puts about RANK_TARGET rows below
  // it, and pass 2 lowers it while the rows below overflow RANK_CAP.
  for (int b = threadIdx.x; b < RANK_BINS; b += blockDim.x)
    hist[b] = 0;
  __syncthreads();
  const int nsample = (R.npool + RANK_SAMPLE - 1) / RANK_SAMPLE;
  for (int i0 = threadIdx.x; i0 < nsample; i0 += UNR * blockDim.x) {
    float d[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int i = i0 + u * blockDim.x;
      const int q = i * RANK_SAMPLE + (int)(((unsigned)i * 2654435761u) >> 29);
      // the load is unconditional (index clamped) so that the UNR loads are in flight together
      const float dq = d32_of(min(q, R.npool - 1));
      d[u] = (i < nsample && q < R.npool) ? dq : -1.0f;
    }
#pragma unroll
    for (int u = 0; u < UNR; u++)
      if (d[u] >= 0.0f)
        atomicAdd(&hist[min(RANK_BINS - 1, (int)(__float_as_uint(d[u]) >> 20))], 1u);
  }
  __syncthreads();
  {
    // the first bin at which the running count reaches RANK_SAMPLE_TARGET (RANK_BINS when it never
    // does): each thread sums PER consecutive bins, a block scan places the sums, and the thread whose
    // range crosses the target finds the bin (the block has 256 threads)
    constexpr int PER = RANK_BINS / 256;
    unsigned loc[PER], sum = 0;
#pragma unroll
    for (int i = 0; i < PER; i++) {
      loc[i] = hist[PER * threadIdx.x + i];
      sum += loc[i];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned v = __shfl_up(incl, off);
      if (lane >= off)
        incl += v;
    }
    if (lane == 63)
      wsum[wv] = incl;
    if (threadIdx.x == 0) {
      all_rows = (R.npool <= RANK_CAP) ? 1 : 0;
      cut_bin = RANK_BINS;
    }
    __syncthreads();
    for (int v = 0; v < wv; v++)
      incl += wsum[v];
    const unsigned excl = incl - sum;
    if (excl < (unsigned)RANK_SAMPLE_TARGET && incl >= (unsigned)RANK_SAMPLE_TARGET) {
      unsigned run = excl;
      int b = PER * threadIdx.x;
#pragma unroll
      for (int i = 0; i < PER; i++) {
        run += loc[i];
        if (run >= (unsigned)RANK_SAMPLE_TARGET) {
          b = PER * threadIdx.x + i;
          break;
        }
      }
      cut_bin = b;
    }
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
    // The pool is in a k-d order, in 64-row chunks with boxes: the chunks of each window of RANK_CAP whose
    // box may hold a row within the cut are listed (in ord, free until the sort), and only their rows are
    // screened.  The squared gap between the target and a chunk's box (outward-rounded FP32 bounds, summed
    // in FP64 and lowered by 2^-40 of itself) bounds the FP64 distance of every row of the chunk from below,
    // so a chunk left out holds no candidate and no zero distance.
    const int nch = (R.npool + 63) / 64;
    for (int cw = 0; cw < nch; cw += RANK_CAP) {
      const int ce = min(nch, cw + RANK_CAP);
      if (threadIdx.x == 0)
        nlist = 0;
      __syncthreads();
      for (int c = cw + (int)threadIdx.x; c < ce; c += blockDim.x) {
        bool keep = true;
        if (!all && R.cbox) {
          const float* bx = R.cbox + (size_t)c * 2 * Ea;
          double lb = 0.0;
          for (int j = 0; j < Ea; j++) {
            const double gp = fmax(0.0, fmax((double)bx[j] - y[j], y[j] - (double)bx[Ea + j]));
            lb += gp * gp;
          }
          keep = lb - lb * 0x1p-40 <= tau;
        }
        if (keep)
          ord[atomicAdd(&nlist, 1)] = c;
      }
      __syncthreads();
      const int nl = nlist;
      // rows of the listed chunks: a wavefront takes one chunk (a row per lane), UNR chunks at a time
      const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
      for (int i0 = 0; i0 < nl; i0 += UNR * 4) {
        float d[UNR];
        int qs[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
          const int li = i0 + 4 * u + wv;
          const int q = li < nl ? ord[li] * 64 + lane : R.npool;
          qs[u] = q;
          const float dq = d32_of(min(q, R.npool - 1)); // unconditional, as in pass 1
          d[u] = q < R.npool ? dq : __builtin_huge_valf();
        }
#pragma unroll
        for (int u = 0; u < UNR; u++) {
          const int q = qs[u];
          if (q >= R.npool || !(d[u] <= T1))
            continue;
          const int row = R.pool[q];
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
      }
      __syncthreads(); // ord is listed again for the next window
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
  // the largest kept distance; drop the group tied with it unless every pool row was examined
  {
    double m = 0.0;
    for (int c = threadIdx.x; c < n; c += blockDim.x)
      m = fmax(m, cs[c]);
    if (threadIdx.x == 0)
      smax_sh = 0.0;
    for (int b = threadIdx.x; b < RANK_NB; b += blockDim.x)
      bcnt[b] = 0u;
    __syncthreads();
    // block max via LDS (values are non-negative doubles: compare as unsigned 64-bit)
    atomicMax((unsigned long long*)&smax_sh, (unsigned long long)__double_as_longlong(m));
    __syncthreads();
  }
  const double smax = smax_sh;
  if (!all_rows)
    for (int c = threadIdx.x; c < n; c += blockDim.x)
      if (cs[c] == smax)
        cs[c] = INF_D;
  // Order by (s, row).  A counting sort into RANK_NB buckets of s over [0, s_max] (the bucket index is
  // monotone in s), then each candidate takes its rank among the members of its bucket; with a bucket
  // of more than RANK_BMAX candidates (many equal distances) the bitonic sort instead.  Rows are
  // distinct, so both give the one order.
  const double scale = (double)RANK_NB / smax;
  const bool bucketable = smax > 0.0 && scale < INF_D;
  auto bucket = [&](double s) { return (int)fmin(s * scale, (double)(RANK_NB - 1)); };
  if (bucketable)
    for (int c = threadIdx.x; c < n; c += blockDim.x)
      if (cs[c] < INF_D)
        atomicAdd(&bcnt[bucket(cs[c])], 1u);
  __syncthreads();
  if (threadIdx.x < 64) {
    // exclusive scan of the bucket counts on wavefront 0 (four buckets a lane), their largest, and the
    // counts cleared for the placement
    static_assert(RANK_NB == 256, "four buckets per lane");
    const int lane = threadIdx.x;
    unsigned v[4], sum = 0u, mx = 0u;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      v[i] = bcnt[4 * lane + i];
      sum += v[i];
      mx = max(mx, v[i]);
    }
    unsigned incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned u = __shfl_up(incl, off);
      if (lane >= off)
        incl += u;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      mx = max(mx, (unsigned)__shfl_xor((int)mx, off));
    unsigned run = incl - sum;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      bstart[4 * lane + i] = run;
      run += v[i];
      bcnt[4 * lane + i] = 0u;
    }
    if (lane == 63)
      bstart[RANK_NB] = incl;
    if (lane == 0)
      bmaxc = mx;
  }
  __syncthreads();
  int kept = 0;
  if (bucketable && bmaxc <= (unsigned)RANK_BMAX) {
    for (int c = threadIdx.x; c < n; c += blockDim.x)
      if (cs[c] < INF_D) {
        const int b = bucket(cs[c]);
        ord[bstart[b] + atomicAdd(&bcnt[b], 1u)] = c;
      }
    __syncthreads();
    kept = min((int)bstart[RANK_NB], RANK_STORE);
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
      const double sc = cs[c];
      if (!(sc < INF_D))
        continue;
      const int qc = cq[c];
      const int b = bucket(sc);
      int r = 0;
      for (int m = (int)bstart[b]; m < (int)bstart[b + 1]; m++) {
        const int o = ord[m];
        const double so = cs[o];
        r += (so < sc || (so == sc && cq[o] < qc)) ? 1 : 0;
      }
      const int p = (int)bstart[b] + r;
      if (p < RANK_STORE) {
        R.rank_row[(size_t)t * R.tstride + (size_t)p * R.jstride] = qc;
        R.rank_s[(size_t)t * R.tstride + (size_t)p * R.jstride] = sc;
      }
    }
  } else {
    int width = 64;
    while (width < n)
      width <<= 1;
    for (int c = n + threadIdx.x; c < width; c += blockDim.x) {
      cs[c] = INF_D;
      cq[c] = INT_MAX;
    }
    __syncthreads();
    // Bitonic sort by (s, row) over the next power of two covering the candidates.  Thread t takes the
    // pairs t, t + 256, ...; for strides up to 64 the pairs of one wavefront stay inside its own
    // 128-element blocks, so only the stages with a longer stride, and their neighbours, need the
    // whole block to wait.
    {
      bool prev_cross = true;
      for (int size = 2; size <= width; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
          const bool cross = stride > 64;
          if (cross || prev_cross) {
            __syncthreads();
          } else {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
          }
          prev_cross = cross;
          for (int t = threadIdx.x; t < width / 2; t += blockDim.x) {
            const int i = ((t & ~(stride - 1)) << 1) | (t & (stride - 1));
            const int j = i + stride;
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
      }
      __syncthreads();
    }
    int kept_b = 0;
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
    kept_b = ncand;
    for (int c = threadIdx.x; c < kept_b; c += blockDim.x) {
      R.rank_row[(size_t)t * R.tstride + (size_t)c * R.jstride] = cq[c];
      R.rank_s[(size_t)t * R.tstride + (size_t)c * R.jstride] = cs[c];
    }
    kept = kept_b;
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
      const int* rr = S.rank_row + (size_t)t * S.tstride;
      const double* rs = S.rank_s + (size_t)t * S.tstride;
      const size_t st = (size_t)S.jstride;
      // four prefix entries per iteration keep several loads in flight
      int j = 0;
      for (; j + 4 <= len && found < k; j += 4) {
        const int r0 = rr[(size_t)j * st], r1 = rr[(size_t)(j + 1) * st], r2 = rr[(size_t)(j + 2) * st],
                  r3 = rr[(size_t)(j + 3) * st];
        const int rows4[4] = { r0, r1, r2, r3 };
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (found < k && inlib(rows4[u])) {
            const double s = rs[(size_t)(j + u) * st];
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
          const double s = rs[(size_t)j * st];
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
          write_unpredicted(S.pred_out, S.rc, S.k_used, ob, S.T, S.np, t, RC_INSUFFICIENT_UNIQUE);
          continue;
        }
      }
      if (ke == 0 || vcnt == 0) {
        write_unpredicted(S.pred_out, S.rc, S.k_used, ob, S.T, S.np, t,
                          S.force_compute ? RC_SUCCESS : RC_INSUFFICIENT_UNIQUE);
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
// One pass per target over its ranked prefix for every rank task at once.  A wavefront owns one
// target; lane l owns the task words l, l + 64, ... of the membership bitmask (WPL words per lane), so
// each task is handled by one lane only.  For prefix entry j (table row r, in (sqrt distance, row)
// order) the lane ANDs row r's membership words with its still-searching tasks; every task hit records
// j, and a task's k-th hit completes it.  Rank tasks have more library rows than k + RANK_ZCAP, so
// k < #valid and the k-NN list is the first k members.  The predictions (src/edm.cpp:707-735) are
// made after the scan, one task per lane at a time.  Tasks still searching when the prefix ends, and
// every task of a target without an exact prefix, go to the exact selection.
__device__ __forceinline__ void hand_over(const ScanParams& S, int task, int t)
{
  const int r = atomicAdd(S.redo_count, 1);
  if (r < S.redo_cap)
    S.redo[r] = make_int2(task, t);
}

// Makes one lane's LDS and global writes visible to the other lanes of its wavefront.
__device__ __forceinline__ void ms_wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}


// S.parts wavefronts share a target, each taking a contiguous range of task words.  Per lane and task
// word the hit counts of its 32 tasks are kept as NP bit planes in registers (bit b of plane p = bit p of
// task b's count), so that counting the hits of one prefix entry and finding the tasks it completes are
// a few bitwise operations; only the prefix position of each hit is stored one by one.  Dynamic LDS per
// wavefront: task words [wp] (unsigned, the still-searching tasks after the scan), then per task KCAP
// prefix positions ([wp * 32 * KCAP] unsigned short), wp = ceil(words / parts).  The predictions go to
// S.stage, target-major, and rank_scatter moves them to the tasks' outputs.
template <int KCAP>
struct Planes
{
  static constexpr int NP = KCAP < 4 ? 2 : KCAP < 8 ? 3 : KCAP < 16 ? 4 : KCAP < 32 ? 5 : 6; // counts 0..KCAP
};

template <int KCAP, int WPL>
__global__ __launch_bounds__(64 * MS_WAVES) void rank_scan_multi(ScanParams S)
{
  constexpr int NP = Planes<KCAP>::NP;
  extern __shared__ unsigned ms_lds[];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gw = blockIdx.x * MS_WAVES + w;
  const int t = gw / S.parts;
  if (t >= S.np)
    return;
  const int wp = (S.words + S.parts - 1) / S.parts;
  const int w0 = (gw % S.parts) * wp;                // first task word of this wavefront
  const int W = max(0, min(S.words - w0, wp));       // its task words
  const int tau_lo = w0 * 32;                        // first task of this wavefront
  const int ntask = max(0, min(S.num_rank - tau_lo, W * 32));
  unsigned* act_sh = ms_lds + (size_t)w * (wp + wp * 16 * KCAP);
  unsigned short* hits = (unsigned short*)(act_sh + wp);

  unsigned act[WPL], cnt[WPL][NP], km[WPL][NP];
  int live = 0;
#pragma unroll
  for (int q = 0; q < WPL; q++) {
    const int wd = lane + 64 * q;
    act[q] = 0u;
#pragma unroll
    for (int p = 0; p < NP; p++) {
      cnt[q][p] = 0u;
      km[q][p] = 0u;
    }
    if (wd < W) {
      const int lo = wd * 32;
      const int nb = min(32, ntask - lo);
      act[q] = nb >= 32 ? ~0u : ((1u << nb) - 1u);
      for (int b = 0; b < nb; b++) {
        const unsigned km1 = (unsigned)(S.k[S.rank_task[tau_lo + lo + b]] - 1); // completing count
#pragma unroll
        for (int p = 0; p < NP; p++)
          km[q][p] |= ((km1 >> p) & 1u) << b;
      }
      live += nb;
    }
  }
  const int len = S.rank_len[t];
  if (S.zlen[t] < 0 || len == 0) {
#pragma unroll
    for (int q = 0; q < WPL; q++)
      for (unsigned m = act[q]; m; m &= m - 1)
        hand_over(S, S.rank_task[tau_lo + (lane + 64 * q) * 32 + __ffs(m) - 1], t);
    return;
  }
  const int* rr = S.rank_row + (size_t)t * S.tstride;
  constexpr int G = 8; // prefix entries whose membership words are loaded together
  for (int c0 = 0; c0 < len; c0 += 64) {
    // stop once every task of every lane has its k neighbours
    if (!__ballot(live > 0))
      break;
    const int n = min(64, len - c0);
    unsigned mg[G][WPL];
    // the prefix rows are contiguous here (jstride 1) and the same for every lane: scalar loads
    auto load_group = [&](int g0) {
      int rows[G];
      if (g0 + G <= n) {
#pragma unroll
        for (int u = 0; u < G; u++)
          rows[u] = rr[c0 + g0 + u];
      } else {
#pragma unroll
        for (int u = 0; u < G; u++)
          rows[u] = g0 + u < n ? rr[c0 + g0 + u] : 0;
      }
#pragma unroll
      for (int u = 0; u < G; u++) {
        const int i = g0 + u;
        const unsigned* mrow = S.member + (size_t)rows[u] * S.words + w0;
#pragma unroll
        for (int q = 0; q < WPL; q++)
          mg[u][q] = (i < n && lane + 64 * q < W) ? mrow[lane + 64 * q] : 0u;
      }
    };
    load_group(0);
    for (int g0 = 0; g0 < n; g0 += G) {
      unsigned cur[G][WPL];
#pragma unroll
      for (int u = 0; u < G; u++)
#pragma unroll
        for (int q = 0; q < WPL; q++)
          cur[u][q] = mg[u][q];
      if (g0 + G < n)
        load_group(g0 + G);
#pragma unroll
      for (int u = 0; u < G; u++) {
        const unsigned short pos = (unsigned short)(c0 + g0 + u);
#pragma unroll
        for (int q = 0; q < WPL; q++) {
          const unsigned m = cur[u][q] & act[q];
          if (!m)
            continue;
          const int base = (lane + 64 * q) * 32;
          // record this entry as hit number count(b) of each task b in m
          for (unsigned mm = m; mm; mm &= mm - 1) {
            const int b = __ffs(mm) - 1;
            int c = 0;
#pragma unroll
            for (int p = 0; p < NP; p++)
              c |= (int)((cnt[q][p] >> b) & 1u) << p;
            hits[(size_t)(base + b) * KCAP + c] = pos;
          }
          // tasks of m whose count was k - 1 are complete; the others count one more hit
          unsigned eq = ~0u;
#pragma unroll
          for (int p = 0; p < NP; p++)
            eq &= ~(cnt[q][p] ^ km[q][p]);
          const unsigned done = m & eq;
          unsigned carry = m;
#pragma unroll
          for (int p = 0; p < NP; p++) {
            const unsigned nc = cnt[q][p] & carry;
            cnt[q][p] ^= carry;
            carry = nc;
          }
          act[q] &= ~done;
          live -= __popc(done);
        }
      }
    }
  }
  // publish the still-searching tasks, then predict the completed ones, U tasks per lane at a time
#pragma unroll
  for (int q = 0; q < WPL; q++)
    if (lane + 64 * q < W)
      act_sh[lane + 64 * q] = act[q];
  ms_wave_sync();
  constexpr int U = KCAP <= 8 ? 2 : 1; // more side by side costs registers, and so occupancy
  const double* rs = S.rank_s + (size_t)t * S.tstride;
  // the weight of entry 0 of a prefix-ordered list (the smallest distance, so a ratio of exactly 1) at the
  // first theta, made once per lane: exp(-theta_0) (src/edm.cpp:707-735)
  const double e0 = fdlibm_exp(-S.thetas[0]);
  for (int tau0 = lane * U; tau0 < ntask; tau0 += 64 * U) {
    int task[U], kq[U];
    bool open_[U];
    int row[U][KCAP];
    double ls[U][KCAP], tg[U][KCAP];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int tau = tau0 + u;
      const bool in = tau < ntask;
      task[u] = in ? S.rank_task[tau_lo + tau] : 0;
      open_[u] = in && ((act_sh[tau >> 5] >> (tau & 31)) & 1u);
      kq[u] = (in && !open_[u]) ? S.k[task[u]] : 0;
#pragma unroll
      for (int j = 0; j < KCAP; j++)
        row[u][j] = j < kq[u] ? hits[(size_t)tau * KCAP + j] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int j = 0; j < KCAP; j++) {
        const size_t e = (size_t)row[u][j] * S.jstride;
        ls[u][j] = j < kq[u] ? rs[e] : 0.0;
        row[u][j] = j < kq[u] ? rr[e] : 0;
      }
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int j = 0; j < KCAP; j++)
        tg[u][j] = j < kq[u] ? S.Ltgt[row[u][j]] : 0.0;
    // hand-overs, then the U predictions side by side (independent exp and divide chains)
    double minDist[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (tau0 + u < ntask && open_[u])
        hand_over(S, task[u], t); // prefix exhausted before k members
      minDist[u] = INF_D;
#pragma unroll
      for (int j = 0; j < KCAP; j++)
        if (j < kq[u])
          minDist[u] = fmin(minDist[u], ls[u][j]);
    }
    for (int th = 0; th < S.T; th++) {
      const double theta = S.thetas[th];
      double w[U][KCAP], sumw[U], r[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        sumw[u] = 0.0;
#pragma unroll
        for (int j = 0; j < KCAP; j++)
          if (j < kq[u]) {
            w[u][j] = j == 0 && th == 0 ? e0 : fdlibm_exp(-theta * (ls[u][j] / minDist[u]));
            sumw[u] = sumw[u] + w[u][j];
          }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        r[u] = 0.0;
#pragma unroll
        for (int j = 0; j < KCAP; j++)
          if (j < kq[u])
            r[u] = r[u] + tg[u][j] * (w[u][j] / sumw[u]);
      }
      double* st = S.stage + ((size_t)th * S.np + t) * S.num_rank + tau_lo;
#pragma unroll
      for (int u = 0; u < U; u++)
        if (tau0 + u < ntask)
          st[tau0 + u] = r[u];
    }
  }
}

// Stage [theta][target][rank task] -> the tasks' outputs (prediction, rc SUCCESS, k used), 64 x 64
// tiles through LDS so that both sides are read and written in whole rows.  Targets handed over to the
// exact selection are overwritten afterwards by it.
__global__ __launch_bounds__(256) void rank_scatter(ScanParams S)
{
  __shared__ double tile[64][65];
  __shared__ int64_t obeg[64]; // output offsets of the tile's 64 tasks
  __shared__ int kk[64];
  const int tau0 = blockIdx.x * 64, t0 = blockIdx.y * 64, th = blockIdx.z;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  if (ty == 0 && tau0 + tx < S.num_rank) {
    const int task = S.rank_task[tau0 + tx];
    obeg[tx] = S.out_begin[task];
    kk[tx] = S.k[task];
  }
  for (int r = ty; r < 64; r += 4) {
    const int t = t0 + r, tau = tau0 + tx;
    tile[r][tx] = (t < S.np && tau < S.num_rank) ? S.stage[((size_t)th * S.np + t) * S.num_rank + tau] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int tau = tau0 + r, t = t0 + tx;
    if (tau < S.num_rank && t < S.np) {
      const int64_t ob = obeg[r];
      S.pred_out[ob + (int64_t)th * S.np + t] = tile[tx][r];
      S.rc[ob + (int64_t)th * S.np + t] = RC_SUCCESS;
      if (th == 0)
        S.k_used[ob / S.T + t] = kk[r];
    }
  }
}

#define FEDM_RANK_M(KC)                                                                                          \
  template __global__ void rank_scan_multi<KC, 1>(ScanParams);                                                   \
  template __global__ void rank_scan_multi<KC, 2>(ScanParams);                                                   \
  template __global__ void rank_scan_multi<KC, 4>(ScanParams);
FEDM_RANK_M(4)
FEDM_RANK_M(8)
FEDM_RANK_M(16)
FEDM_RANK_M(32)

#define FEDM_RANK_S(KC, EM) template __global__ void rank_scan<KC, EM>(ScanParams);
#define FEDM_RANK_SE(KC) FEDM_RANK_S(KC, 4) FEDM_RANK_S(KC, 8) FEDM_RANK_S(KC, 16) FEDM_RANK_S(KC, 32)
FEDM_RANK_SE(4)
FEDM_RANK_SE(8)
FEDM_RANK_SE(16)
FEDM_RANK_SE(32)

} // namespace fedm
