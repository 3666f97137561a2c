// Exact k-nearest-neighbour selection for k wider than the register lists (64 < k < #valid).
//
// Runs on the dense route after dense_dist_f64 has written one row of distances per target
// (valid entries >= 0, invalid -1).  The reference takes the first k entries of the ascending
// (distance, library position) order (src/edm.cpp:652-681, stable_sort or partial_sort with an index
// tiebreak, which agree).  Non-negative doubles order like their bit patterns, so one block per target
// runs an MSB-first radix select on the 64-bit pattern to find the k-th smallest distance K and how many
// entries equal to K it needs; a position-ordered compaction takes every entry below K and the first
// of those equal to K; a bitonic sort on (pattern, position) then gives the neighbour order.  The
// lists have the layout finish_lists writes, so the S-map kernel reads them unchanged; the simplex
// prediction (src/edm.cpp:707-735) is made here in neighbour order.
#include "kernels_common.h"
#include "fdlibm_exp.h"

namespace fedm {

constexpr int SW_THREADS = 256;
constexpr int SW_BINS = 4096;      // 12-bit digits
constexpr int SW_LDS_SORT = 4096;  // lists up to this (padded) size are sorted in LDS

__device__ __forceinline__ bool kv_less(unsigned long long a, int ia, unsigned long long b, int ib)
{
  return a < b || (a == b && ia < ib);
}

// Exclusive prefix sum of one value per thread over a 256-thread block; *total gets the block sum.
__device__ __forceinline__ int block_excl_scan(int v, int* wsum, int* total)
{
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o)
      x += y;
  }
  if (lane == 63)
    wsum[w] = x;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < SW_THREADS / 64; i++) {
    off += i < w ? wsum[i] : 0;
    tot += wsum[i];
  }
  __syncthreads();
  *total = tot;
  return off + x - v;
}

// Bitonic sort of n2 (a power of two) (key, index) pairs, ascending.
__device__ void bitonic_sort(unsigned long long* key, int* idx, int n2)
{
  for (int size = 2; size <= n2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < (n2 >> 1); i += SW_THREADS) {
        const int lo = 2 * stride * (i / stride) + (i % stride);
        const int hi = lo + stride;
        const bool asc = (lo & size) == 0;
        const unsigned long long a = key[lo], b = key[hi];
        const int ia = idx[lo], ib = idx[hi];
        if (kv_less(b, ib, a, ia) == asc) {
          key[lo] = b;
          key[hi] = a;
          idx[lo] = ib;
          idx[hi] = ia;
        }
      }
      __syncthreads();
    }
  }
}

// One block per target t0 + blockIdx.x of the dense block; only targets whose policy step left
// nn_k[t] = k > 0 are selected.  Lists go to nn_i/nn_s at row (t - t0) with stride F.nn_stride.
// gkey/gidx: [nt x gstride] scratch for lists whose padded size exceeds SW_LDS_SORT.
__global__ __launch_bounds__(SW_THREADS) void select_wide(FinishParams F, const double* __restrict__ D, int t0, int nt,
                                                          unsigned long long* __restrict__ gkey, int* __restrict__ gidx,
                                                          int gstride)
{
  __shared__ union
  {
    unsigned int hist[SW_BINS];
    unsigned long long key[SW_LDS_SORT];
  } sh;
  __shared__ int sidx[SW_LDS_SORT];
  __shared__ int wsum[SW_THREADS / 64];
  __shared__ int s_digit, s_below;

  const int tl = blockIdx.x;
  const int t = t0 + tl;
  if (tl >= nt || t >= F.NP)
    return;
  const int k = F.nn_k[t];
  if (k <= 0)
    return;
  const int NL = F.NL;
  const double* row = D + (size_t)tl * NL;
  const int tid = threadIdx.x;

  // 1. radix select of the k-th smallest bit pattern among the valid entries
  unsigned long long prefix = 0;
  int kr = k;
  // digits: bits 63..52, 51..40, 39..28, 27..16, 15..4, 3..0
  for (int p = 0; p < 6; p++) {
    const int sft = p < 5 ? 52 - 12 * p : 0, nb = p < 5 ? 12 : 4, top = sft + nb;
    const int nbins = 1 << nb;
    for (int b = tid; b < nbins; b += SW_THREADS)
      sh.hist[b] = 0;
    __syncthreads();
    for (int i = tid; i < NL; i += SW_THREADS) {
      const double d = row[i];
      if (!(d >= 0.0))
        continue;
      const unsigned long long key = (unsigned long long)__double_as_longlong(d);
      if (top < 64 && (key >> top) != (prefix >> top))
        continue;
      atomicAdd(&sh.hist[(key >> sft) & (unsigned long long)(nbins - 1)], 1u);
    }
    __syncthreads();
    const int per = (nbins + SW_THREADS - 1) / SW_THREADS;
    const int b0 = tid * per;
    int local = 0;
    for (int b = b0; b < min(nbins, b0 + per); b++)
      local += (int)sh.hist[b];
    int total;
    const int excl = block_excl_scan(local, wsum, &total);
    if (local > 0 && excl < kr && kr <= excl + local) {
      int c = excl;
      for (int b = b0; b < min(nbins, b0 + per); b++) {
        const int h = (int)sh.hist[b];
        if (c + h >= kr) {
          s_digit = b;
          s_below = c;
          break;
        }
        c += h;
      }
    }
    __syncthreads();
    prefix |= (unsigned long long)s_digit << sft;
    kr -= s_below;
    __syncthreads();
  }
  const unsigned long long kth = prefix;

  // 2. position-ordered compaction: every entry below K, then the first kr entries equal to K
  const int n2 = k <= 1 ? 1 : 1 << (32 - __clz(k - 1));
  unsigned long long* key = n2 <= SW_LDS_SORT ? sh.key : gkey + (size_t)tl * gstride;
  int* idx = n2 <= SW_LDS_SORT ? sidx : gidx + (size_t)tl * gstride;
  int base_sel = 0, base_eq = 0;
  for (int c0 = 0; c0 < NL && base_sel < k; c0 += SW_THREADS) {
    const int i = c0 + tid;
    const double d = i < NL ? row[i] : -1.0;
    const bool valid = d >= 0.0;
    const unsigned long long kb = (unsigned long long)__double_as_longlong(d);
    const bool lt = valid && kb < kth;
    const bool eq = valid && kb == kth;
    int eq_tot, sel_tot;
    const int eq_excl = block_excl_scan(eq ? 1 : 0, wsum, &eq_tot);
    const bool take = lt || (eq && base_eq + eq_excl < kr);
    const int sel_excl = block_excl_scan(take ? 1 : 0, wsum, &sel_tot);
    if (take && base_sel + sel_excl < k) {
      key[base_sel + sel_excl] = kb;
      idx[base_sel + sel_excl] = i;
    }
    base_eq += eq_tot;
    base_sel += sel_tot;
  }
  for (int j = k + tid; j < n2; j += SW_THREADS) {
    key[j] = ~0ull;
    idx[j] = INT_MAX;
  }
  __syncthreads();

  // 3. neighbour order
  bitonic_sort(key, idx, n2);

  int* oi = F.nn_i + (size_t)tl * F.nn_stride;
  double* os = F.nn_s + (size_t)tl * F.nn_stride;
  for (int j = tid; j < k; j += SW_THREADS) {
    oi[j] = idx[j];
    os[j] = __longlong_as_double((long long)key[j]);
  }
  if (F.nn_out)
    for (int j = tid; j < F.nn_cap; j += SW_THREADS)
      F.nn_out[(size_t)t * F.nn_cap + j] = j < k ? idx[j] : -1;

  // 4. simplex, summed in neighbour order as the CPU path does
  if (F.algorithm == ALG_SIMPLEX && tid == 0) {
    double minDist = __longlong_as_double((long long)key[0]);
    for (int j = 1; j < k; j++)
      minDist = fmin(minDist, __longlong_as_double((long long)key[j]));
    for (int th = 0; th < F.T; th++) {
      const double theta = F.thetas[th];
      double sumw = 0.0;
      for (int j = 0; j < k; j++)
        sumw = sumw + fdlibm_exp(-theta * (__longlong_as_double((long long)key[j]) / minDist));
      double r = 0.0;
      for (int j = 0; j < k; j++)
        r = r + F.targets[idx[j]] * (fdlibm_exp(-theta * (__longlong_as_double((long long)key[j]) / minDist)) / sumw);
      F.pred[(size_t)th * F.NP + t] = r;
      F.rc[(size_t)th * F.NP + t] = RC_SUCCESS;
    }
  }
}

} // namespace fedm
