// Batched simplex prediction over many tasks that share one point table (CCM library-size x
// replicate sweeps, E/theta grids): one persistent launch drains a work queue of (task, block of 256
// targets) items, heaviest libraries first.
//
// The library points are gathered through the task's row list into LDS in FP64, zero padded to EMAX
// coordinates (a zero coordinate adds exactly 0 to the sum, so the padded sum is the CPU path's sum,
// src/distances.cpp:186-216).  Both precision forms use this kernel: FP64 arithmetic runs at the FP32
// rate on CDNA4 vector units, so FP32 screening saves nothing here.  The cost lies in the list
// updates, which diverge across a wavefront: each thread screens a block of rows (a library's first 16,
// then 32 at a time) against its threshold first and then walks only the rows that passed, so a
// wavefront runs as many updates as its busiest thread needs rather than one per row in which any
// thread updates.  The screen after the first block is an expanded FP64 form with a rounding bound
// (three fused multiply-adds a row); the walk recomputes each passing row's distance in coordinate
// order.
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
// after equal squared distances, and every entry behind it moves one place down.  Every compare is made
// first: the list is sorted, so the compares are monotone, entry j takes entry j - 1 when the new one goes
// before j - 1, the new one when it goes at j, and keeps its own otherwise.  Updated from the last entry
// down, each entry is overwritten in place, which spares the register copies a chained insertion costs
// in a loop.  Returns the squared distance that leaves the list.
template <int KCAP>
__device__ __forceinline__ double insert_acc(double (&la)[KCAP], int (&li)[KCAP], double a, int i)
{
  bool c[KCAP];
#pragma unroll
  for (int j = 0; j < KCAP; j++)
    c[j] = a < la[j];
  const double out = c[KCAP - 1] ? la[KCAP - 1] : a;
#pragma unroll
  for (int j = KCAP - 1; j >= 1; j--) {
    const double v = c[j] ? a : la[j];
    const int vi = c[j] ? i : li[j];
    la[j] = c[j - 1] ? la[j - 1] : v;
    li[j] = c[j - 1] ? li[j - 1] : vi;
  }
  la[0] = c[0] ? a : la[0];
  li[0] = c[0] ? i : li[0];
  return out;
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
// tgt: the task's library targets gathered by position, or null to read them through the row list.
// eth: theta_exps of the call in LDS, or null: the weight of the first entry while the list stays in
// distance order.
template <int KCAP>
__device__ void finish_target(const BatchParams& B, int tk, int t, int np, const int* lrows, int nl, int zeros,
                              int k, double (&ls)[KCAP], int (&li)[KCAP], const double* tgt = nullptr,
                              const double* eth = nullptr)
{
  const int cnt = nl - zeros;
  const int T = B.T;
  const int64_t ob = B.out_begin[tk];
  int ke = k;
  if (ke > cnt) {
    if (B.force_compute) {
      ke = cnt;
    } else {
      write_unpredicted(B.pred, B.rc, B.k_used, ob, T, np, t, RC_INSUFFICIENT_UNIQUE);
      return;
    }
  }
  if (ke == 0 || cnt == 0) {
    write_unpredicted(B.pred, B.rc, B.k_used, ob, T, np, t, B.force_compute ? RC_SUCCESS : RC_INSUFFICIENT_UNIQUE);
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
    tg[j] = j < ke ? (tgt ? tgt[li[j]] : B.Ltgt[lrows[li[j]]]) : 0.0;
  // entry 0 holds the smallest distance unless the list was put in library order
  const bool first_min = eth && ke != cnt;
  for (int th = 0; th < T; th++) {
    const double theta = B.thetas[th];
    double w0;
    if (first_min && th < ETH_N)
      w0 = eth[th];
    else
      w0 = fdlibm_exp(-theta * (ls[0] / minDist));
    double w[KCAP];
    double sumw = 0.0;
#pragma unroll
    for (int j = 0; j < KCAP; j++)
      if (j < ke) {
        w[j] = j == 0 ? w0 : fdlibm_exp(-theta * (ls[j] / minDist));
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

// FP64 squared distance of one LDS row (EMAX coordinates, zero padded) to y, in coordinate order.
template <int EMAX>
__device__ __forceinline__ double row_dist(const double* x, const double (&y)[EMAX])
{
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < EMAX; j++) {
    const double d = x[j] - y[j];
    acc += d * d;
  }
  return acc;
}

constexpr int SUB0 = SUB0_ROWS; // rows of a library's first block, screened against their own k-th distance
constexpr int SUB = 32;  // rows screened together against the threshold they start with (one mask bit each)

// LDS row of the direct kernel: EMAX coordinates (zero padded), the squared norm, padded to an even
// number of doubles so that rows stay 16-byte aligned
template <int EMAX>
constexpr int row_stride() { return (EMAX + 2) & ~1; }

// Appends a row's screen result to a block mask: after a block's rows, row r of the block is bit
// 31 - r when the block has 32 rows.  The empty asm keeps the compiler from holding every row's bit in a
// register of its own until the block ends.
__device__ __forceinline__ unsigned shift_in(unsigned mask, bool pass)
{
  mask = (mask << 1) | (pass ? 1u : 0u);
  asm volatile("" : "+v"(mask));
  return mask;
}

// The rows of `mask` (bit 31 - r: LDS row x + r ST, library position pos0 + r) in position order, each
// against the list as it stands.  The update runs on every lane without branches: a row that does not
// enter passes an infinite key, which changes nothing.
template <int KCAP, int EMAX, bool FULL>
__device__ __forceinline__ void walk_rows(unsigned mask, const double* x, const double (&y)[EMAX], int pos0, int k,
                                          double (&la)[KCAP], int (&li)[KCAP], double& kth, double& thr_g, int& zeros,
                                          bool& near)
{
  constexpr int ST = row_stride<EMAX>();
  while (mask) {
    const int r = __builtin_clz(mask);
    mask ^= 0x80000000u >> r;
    const double acc = row_dist<EMAX>(x + r * ST, y);
    // keep rule; a sum equal to MISSING cannot occur on this route (coordinates below 1e30)
    const bool drop = acc == 0.0 || acc == MISSING_D;
    zeros += drop ? 1 : 0;
    if (FULL) {
      // k = KCAP: the squared distance that leaves the list is the row's own when it does not enter
      // and the old k-th when it does; either lies strictly between the new k-th and the k-th times
      // (1 + 2^-49) when its square root may equal the k-th's
      const double out = insert_acc<KCAP>(la, li, drop ? INF_D : acc, pos0 + r);
      kth = la[KCAP - 1];
      near = near | ((out > kth) & (out < kth * SQRT_GUARD));
    } else {
      // kth < acc < kth (1 + 2^-49): its square root may equal the k-th's
      near = near | (!drop & (acc > kth) & (acc < thr_g));
      const double out = kth; // leaves the k-th place when the row enters
      insert_acc<KCAP>(la, li, (!drop & (acc < kth)) ? acc : INF_D, pos0 + r);
      kth = kth_of<KCAP>(la, k);
      near = near | ((out > kth) & (out < kth * SQRT_GUARD));
    }
    thr_g = kth * SQRT_GUARD;
  }
}

// Waves per SIMD the short-list instances are held to (batch_waves, batch_common.h): at six (80 VGPRs) a
// few registers spill, and the kernel still runs faster than at five (94 VGPRs, no spills), being bound
// by issue more than by latency.  WPE is the waves per SIMD of an instance; the autotuner can pick the
// batch_waves_alt instance instead.
template <int KCAP, int EMAX, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void batch_simplex(BatchParams B)
{
  // library rows per LDS chunk (a multiple of SUB); rows past the chunk's end hold +inf and never pass
  constexpr int ST = row_stride<EMAX>();
  constexpr int BCH = EMAX <= 4 ? 512 : (4096 / ST) / SUB * SUB;
  // the first chunk's blocks start at SUB0, SUB0 + SUB, ...: its last block may reach SUB0 rows past BCH
  __shared__ double lds64[(BCH + SUB0) * ST];
  __shared__ int item_sh;
  __shared__ double eth[ETH_N];

  const int Ea = B.Ea;
  theta_exps(B.thetas, B.T, eth);
  if (threadIdx.x == 0)
    item_sh = atomicAdd(B.counter, 1);
  __syncthreads();
  int item = item_sh;
  while (item < B.num_items) {
    __syncthreads();
    // the next item is claimed now, so that the claim's latency overlaps this item's work
    if (threadIdx.x == 0)
      item_sh = atomicAdd(B.counter, 1);
    do {
      const int2 it = B.items[item];
      const int tk = it.x;
      const int np = B.pred_count[tk];
      const int nl = B.lib_count[tk];
      const int t = it.y + threadIdx.x;
      const bool active = t < np;
      const int* lrows = B.lib_rows + B.lib_begin[tk];
      const int k = B.k[tk];
      // this task's library rows and targets, gathered contiguously by position
      const double* lg = B.Lg + (size_t)B.gbeg[tk] * Ea;
      const double* tgt = B.Tg + B.gbeg[tk];

      double y[EMAX];
      double yy = 0.0;
      const double* yp = B.Pg + (size_t)(B.pred_begin[tk] + (active ? t : 0)) * Ea;
#pragma unroll
      for (int j = 0; j < EMAX; j++) {
        y[j] = (active && j < Ea) ? yp[j] : 0.0;
        yy += y[j] * y[j];
      }
      double la[KCAP];
      int li[KCAP];
#pragma unroll
      for (int j = 0; j < KCAP; j++) {
        la[j] = INF_D;
        li[j] = INT_MAX;
      }
      double kth = INF_D, thr_g = INF_D;
      int zeros = 0;
      bool near = false;

      for (int c0 = 0; c0 < nl; c0 += BCH) {
        const int cn = min(BCH, nl - c0);
        const int cpad = c0 == 0 ? (cn <= SUB0 ? SUB0 : SUB0 + (cn - SUB0 + SUB - 1) / SUB * SUB)
                                 : (cn + SUB - 1) / SUB * SUB;
        __syncthreads();
        for (int c = threadIdx.x; c < cpad; c += blockDim.x) {
          double* dst = lds64 + c * ST;
          if (c < cn) {
            const double* src = lg + (size_t)(c0 + c) * Ea;
            double xx = 0.0;
#pragma unroll
            for (int j = 0; j < EMAX; j++) {
              const double v = j < Ea ? src[j] : 0.0;
              dst[j] = v;
              xx += v * v;
            }
            dst[EMAX] = -0.5 * xx;
          } else {
            // infinite distance, and an expanded screen value of -inf or NaN, which never passes
#pragma unroll
            for (int j = 0; j < EMAX; j++)
              dst[j] = j == 0 ? INF_D : 0.0;
            dst[EMAX] = -INF_D;
          }
        }
        __syncthreads();
        if (!active)
          continue;
        for (int s0 = 0; s0 < cpad;) {
          // Screen a block of rows against a threshold, then walk the rows that pass in position order
          // with the list as it stands.  Any threshold at or above (final k-th) (1 + 2^-49) gives the
          // same list, zero count and hand-over decision as walking every row: the rows it drops can
          // neither enter the final list nor lie just above its k-th entry.  The current threshold is
          // one (it only falls); for a library's first SUB0 rows, before it is finite, the KCAP-th
          // smallest nonzero distance among them is another.  Rows at distance 0 always pass (every
          // bound is > 0).
          unsigned mask = 0;
          int nb;
          if (c0 == 0 && s0 == 0) {
            nb = SUB0;
            if (KCAP * 4 <= SUB0) {
              double dr[SUB0], m[KCAP];
#pragma unroll
              for (int j = 0; j < KCAP; j++)
                m[j] = INF_D;
#pragma unroll
              for (int r = 0; r < SUB0; r++) {
                dr[r] = row_dist<EMAX>(lds64 + r * ST, y);
                double v = dr[r] == 0.0 ? INF_D : dr[r];
#pragma unroll
                for (int j = 0; j < KCAP; j++) {
                  const double lo = fmin(m[j], v);
                  v = fmax(m[j], v);
                  m[j] = lo;
                }
              }
              const double t0 = m[KCAP - 1] == INF_D ? INF_D : m[KCAP - 1] * SQRT_GUARD;
#pragma unroll
              for (int r = 0; r < SUB0; r++)
                mask = shift_in(mask, dr[r] < t0);
            } else {
#pragma unroll
              for (int r = 0; r < SUB0; r++)
                mask = shift_in(mask, row_dist<EMAX>(lds64 + r * ST, y) < thr_g);
            }
            mask <<= 32 - SUB0;
          } else {
            // Expanded screen: e = x.y - |x|^2 / 2 by fused multiply-adds, so that the distance is
            // |y|^2 - 2 e; the row passes when |y|^2 - 2 e < thr_g + c4, c4 bounding the rounding of
            // both sides (the direct sum's included).  Every row whose distance is below thr_g
            // passes; a row that passes needlessly lies at or above the k-th distance times
            // (1 + 2^-49) and changes nothing in the walk.
            nb = SUB;
            const double lo = ((yy - thr_g) - B.c4) * 0.5;
#pragma unroll
            for (int r = 0; r < SUB; r++) {
              const double* x = lds64 + (s0 + r) * ST;
              double e = x[EMAX];
#pragma unroll
              for (int j = 0; j < EMAX; j++)
                e = __builtin_fma(x[j], y[j], e);
              mask = shift_in(mask, e > lo);
            }
          }
          // the passing rows in position order, each against the list as it stands (k = KCAP, the
          // common case, reads the k-th entry directly)
          if (k == KCAP)
            walk_rows<KCAP, EMAX, true>(mask, lds64 + s0 * ST, y, c0 + s0, k, la, li, kth, thr_g, zeros, near);
          else
            walk_rows<KCAP, EMAX, false>(mask, lds64 + s0 * ST, y, c0 + s0, k, la, li, kth, thr_g, zeros, near);
          s0 += nb;
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
      finish_target<KCAP>(B, tk, t, np, lrows, nl, zeros, k, ls, li, tgt, eth);
    } while (0);
    __syncthreads();
    item = item_sh;
  }
}

// ---------------------------------------------------------------------------------------------------
// The direct kernel with spatial pruning, for libraries of PR_MIN..PR_BCH rows and points of at most
// PR_EMAX coordinates.  The 256 targets of an item are consecutive in a spatial (Morton) order of the
// prediction list, so they lie close together.  The library rows are staged in PR_NB buckets of
// increasing distance r from the targets' centroid c (the bucket index is monotone in the computed r,
// and each bucket keeps the smallest r it holds).  Before each block of rows a lane tests the triangle
// inequality against the smallest r still to come: |x - y| >= r - |y - c|.  With the FP64 rounding of
// both distances bounded (relative (E + 3) 2^-53 each; the test takes (E + 6) 2^-52 of their sum) and a margin that covers the
// rounding of the coordinate-order squared distance, a lane whose bound exceeds sqrt(threshold) can
// stop: no remaining row can enter its list, be dropped as a zero distance, or lie just above its k-th
// entry.  A wavefront leaves the block loop when all its lanes can.  Rows no longer arrive in position
// order, so list entries are compared by (squared distance, position); screening, walks, hand-over and
// finish are the direct kernel's.

constexpr int PR_NB = 64;
// waves per SIMD of the pruned kernel: four (128 VGPRs, 3 spilled) measured 15 % faster than the
// unconstrained three (147 VGPRs)
constexpr int PR_WAVES = 4;

template <int KCAP>
__device__ __forceinline__ double insert_lex(double (&la)[KCAP], int (&li)[KCAP], double a, int i)
{
  bool c[KCAP];
#pragma unroll
  for (int j = 0; j < KCAP; j++)
    c[j] = a < la[j] || (a == la[j] && i < li[j]);
  const double out = c[KCAP - 1] ? la[KCAP - 1] : a;
#pragma unroll
  for (int j = KCAP - 1; j >= 1; j--) {
    const double v = c[j] ? a : la[j];
    const int vi = c[j] ? i : li[j];
    la[j] = c[j - 1] ? la[j - 1] : v;
    li[j] = c[j - 1] ? li[j - 1] : vi;
  }
  la[0] = c[0] ? a : la[0];
  li[0] = c[0] ? i : li[0];
  return out;
}

template <int KCAP>
__device__ __forceinline__ int kth_pos(const int (&li)[KCAP], int k)
{
  int v = INT_MAX;
#pragma unroll
  for (int j = 0; j < KCAP; j++)
    if (j == k - 1)
      v = li[j];
  return v;
}

// walk_rows with each row's library position read from pos[] (rows in bucket order).
template <int KCAP, int EMAX, bool FULL>
__device__ __forceinline__ void walk_rows_pos(unsigned mask, const double* x, const int* pos, const double (&y)[EMAX],
                                              int k, double (&la)[KCAP], int (&li)[KCAP], double& kth, double& thr_g,
                                              int& zeros, bool& near)
{
  constexpr int ST = row_stride<EMAX>();
  while (mask) {
    const int r = __builtin_clz(mask);
    mask ^= 0x80000000u >> r;
    const double acc = row_dist<EMAX>(x + r * ST, y);
    const int pr = pos[r];
    const bool drop = acc == 0.0 || acc == MISSING_D;
    zeros += drop ? 1 : 0;
    // a row that does not enter passes (inf, INT_MAX), which compares equal to an empty entry and so
    // changes nothing
    if (FULL) {
      const double out = insert_lex<KCAP>(la, li, drop ? INF_D : acc, drop ? INT_MAX : pr);
      kth = la[KCAP - 1];
      near = near | ((out > kth) & (out < kth * SQRT_GUARD));
    } else {
      // rows arrive out of position order: one equal to the k-th entry enters when its position is lower
      const int ikth = kth_pos<KCAP>(li, k);
      const bool enters = !drop & ((acc < kth) | ((acc == kth) & (pr < ikth)));
      near = near | (!drop & (acc > kth) & (acc < thr_g));
      const double out = kth;
      insert_lex<KCAP>(la, li, enters ? acc : INF_D, enters ? pr : INT_MAX);
      kth = kth_of<KCAP>(la, k);
      near = near | ((out > kth) & (out < kth * SQRT_GUARD));
    }
    thr_g = kth * SQRT_GUARD;
  }
}

template <int KCAP, int EMAX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PR_WAVES))) void batch_simplex_pruned(BatchParams B)
{
  constexpr int ST = row_stride<EMAX>();
  constexpr int NBLK = 2 + (PR_BCH - SUB0 + SUB - 1) / SUB; // row blocks of a library, plus one
  __shared__ double lds64[(PR_BCH + SUB0) * ST];
  __shared__ int lpos[PR_BCH + SUB0];
  __shared__ double rr[PR_BCH];
  __shared__ int bk[PR_BCH], rk[PR_BCH], sb[PR_BCH];
  __shared__ int bstart[PR_NB + 1];
  __shared__ unsigned long long bmin[PR_NB];
  __shared__ double blo[NBLK];
  __shared__ double wsum[4][EMAX];
  __shared__ unsigned long long rext[2];
  __shared__ int item_sh;
  __shared__ double eth[ETH_N];

  const int Ea = B.Ea;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  theta_exps(B.thetas, B.T, eth);
  // slack factors of the pruning test (see above)
  const double eta2 = (Ea + 6) * 2.220446049250313e-16, eta3 = (8.0 * Ea + 16.0) * 1.1102230246251565e-16;
  // the next item: from this block's XCD queue, then from the others in turn (-1 when all are empty)
  const int nq = B.pq_n;
  const int home = blockIdx.x % nq;
  auto claim = [&B, home, nq]() -> int {
    for (int d = 0; d < nq; d++) {
      const int q = (home + d) % nq;
      const int n = B.pq_begin[q + 1] - B.pq_begin[q];
      if (n == 0)
        continue;
      const int i = atomicAdd(&B.pcounter[q * PR_CSTRIDE], 1);
      if (i < n)
        return B.pq_begin[q] + i;
    }
    return -1;
  };
  if (tid == 0)
    item_sh = claim();
  __syncthreads();
  int item = item_sh;
  while (item >= 0) {
    __syncthreads();
    if (tid == 0)
      item_sh = claim();
    do {
      const int2 it = B.pitems[item];
      const int tk = it.x;
      const int np = B.pred_count[tk];
      const int nl = B.lib_count[tk];
      const int ts = it.y + tid;
      const bool active = ts < np;
      const int t = active ? B.tperm[B.pred_begin[tk] + ts] : 0;
      const int* lrows = B.lib_rows + B.lib_begin[tk];
      const int k = B.k[tk];
      const double* lg = B.Lg + (size_t)B.gbeg[tk] * Ea;
      const double* tgt = B.Tg + B.gbeg[tk];

      double y[EMAX];
      const double* yp = B.Pg + (size_t)(B.pred_begin[tk] + t) * Ea;
#pragma unroll
      for (int j = 0; j < EMAX; j++)
        y[j] = (active && j < Ea) ? yp[j] : 0.0;
      double yy = 0.0;
#pragma unroll
      for (int j = 0; j < EMAX; j++)
        yy += y[j] * y[j];
      // centroid of the item's targets (any point serves; only the rounding of the distances matters)
#pragma unroll
      for (int j = 0; j < EMAX; j++) {
        double v = y[j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
          v += __shfl_xor(v, o);
        if (lane == 0)
          wsum[w][j] = v;
      }
      if (tid == 0) {
        rext[0] = ~0ull;
        rext[1] = 0ull;
      }
      for (int b = tid; b < PR_NB; b += blockDim.x) {
        bstart[b] = 0;
        bmin[b] = ~0ull;
      }
      __syncthreads();
      const double na = (double)min(256, np - it.y);
      double cen[EMAX];
#pragma unroll
      for (int j = 0; j < EMAX; j++)
        cen[j] = ((wsum[0][j] + wsum[1][j]) + (wsum[2][j] + wsum[3][j])) / na;
      double rho2 = 0.0;
#pragma unroll
      for (int j = 0; j < EMAX; j++) {
        const double d = y[j] - cen[j];
        rho2 += d * d;
      }
      const double rho = sqrt(rho2);
      // distances of the library rows from the centroid, and their range
      for (int c = tid; c < nl; c += blockDim.x) {
        const double* src = lg + (size_t)c * Ea;
        double a = 0.0;
#pragma unroll
        for (int j = 0; j < EMAX; j++) {
          const double d = (j < Ea ? src[j] : 0.0) - cen[j];
          a += d * d;
        }
        rr[c] = sqrt(a);
      }
      {
        // the range of r: per-thread, then per-wavefront, then one atomic per wavefront
        double mn = INF_D, mx = 0.0;
        for (int c = tid; c < nl; c += blockDim.x) {
          mn = fmin(mn, rr[c]);
          mx = fmax(mx, rr[c]);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          mn = fmin(mn, __shfl_xor(mn, o));
          mx = fmax(mx, __shfl_xor(mx, o));
        }
        if (lane == 0) {
          atomicMin(&rext[0], (unsigned long long)__double_as_longlong(mn));
          atomicMax(&rext[1], (unsigned long long)__double_as_longlong(mx));
        }
      }
      __syncthreads();
      const double rmin = __longlong_as_double((long long)rext[0]);
      const double rmax = __longlong_as_double((long long)rext[1]);
      const double width = (rmax - rmin) / (double)PR_NB;
      for (int c = tid; c < nl; c += blockDim.x) {
        const double r = rr[c];
        int b = 0;
        if (width > 0.0) {
          const double q = (r - rmin) / width;
          b = q >= (double)(PR_NB - 1) ? PR_NB - 1 : (int)q;
        }
        bk[c] = b;
        rk[c] = atomicAdd(&bstart[b], 1);
        atomicMin(&bmin[b], (unsigned long long)__double_as_longlong(r));
      }
      __syncthreads();
      if (w == 0) {
        // exclusive scan of the PR_NB (= 64) bucket counts on one wavefront
        const int cnt = bstart[lane];
        int v = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int u = __shfl_up(v, o);
          if (lane >= o)
            v += u;
        }
        bstart[lane] = v - cnt;
        if (lane == 63)
          bstart[PR_NB] = v;
      }
      __syncthreads();
      const int cpad = SUB0 + (nl - SUB0 + SUB - 1) / SUB * SUB;
      for (int c = tid; c < cpad; c += blockDim.x) {
        if (c < nl) {
          const int slot = bstart[bk[c]] + rk[c];
          double* dst = lds64 + slot * ST;
          const double* src = lg + (size_t)c * Ea;
          double xx = 0.0;
#pragma unroll
          for (int j = 0; j < EMAX; j++) {
            const double v = j < Ea ? src[j] : 0.0;
            dst[j] = v;
            xx += v * v;
          }
          dst[EMAX] = -0.5 * xx;
          lpos[slot] = c;
          sb[slot] = bk[c];
        } else {
          double* dst = lds64 + c * ST;
#pragma unroll
          for (int j = 0; j < EMAX; j++)
            dst[j] = j == 0 ? INF_D : 0.0;
          dst[EMAX] = -INF_D;
          lpos[c] = INT_MAX;
        }
      }
      __syncthreads();
      // smallest centroid distance among the rows from each block start on: the bmin of the bucket that
      // holds the block's first slot (buckets are in increasing r)
      for (int q = tid; q < NBLK; q += blockDim.x) {
        const int s0 = q == 0 ? 0 : SUB0 + (q - 1) * SUB;
        blo[q] = s0 < nl ? __longlong_as_double((long long)bmin[sb[s0]]) : INF_D;
      }
      __syncthreads();
      if (!active)
        continue;

      double la[KCAP];
      int li[KCAP];
#pragma unroll
      for (int j = 0; j < KCAP; j++) {
        la[j] = INF_D;
        li[j] = INT_MAX;
      }
      double kth = INF_D, thr_g = INF_D;
      int zeros = 0;
      bool near = false;
      for (int s0 = 0, q = 0; s0 < cpad; q++) {
        unsigned mask = 0;
        int nb;
        if (s0 == 0) {
          nb = SUB0;
          if (KCAP * 4 <= SUB0) {
            double dr[SUB0], m[KCAP];
#pragma unroll
            for (int j = 0; j < KCAP; j++)
              m[j] = INF_D;
#pragma unroll
            for (int r = 0; r < SUB0; r++) {
              dr[r] = row_dist<EMAX>(lds64 + r * ST, y);
              double v = dr[r] == 0.0 ? INF_D : dr[r];
#pragma unroll
              for (int j = 0; j < KCAP; j++) {
                const double lo = fmin(m[j], v);
                v = fmax(m[j], v);
                m[j] = lo;
              }
            }
            const double t0 = m[KCAP - 1] == INF_D ? INF_D : m[KCAP - 1] * SQRT_GUARD;
#pragma unroll
            for (int r = 0; r < SUB0; r++)
              mask = shift_in(mask, dr[r] < t0);
          } else {
#pragma unroll
            for (int r = 0; r < SUB0; r++)
              mask = shift_in(mask, row_dist<EMAX>(lds64 + r * ST, y) < thr_g);
          }
          mask <<= 32 - SUB0;
        } else {
          // the triangle-inequality exit: no row from here on can matter to this lane
          const double lo = blo[q];
          const bool done = lo == INF_D || lo - rho - (lo + rho) * eta2 > sqrt(thr_g) * (1.0 + eta3);
          if (__ballot(!done) == 0ull)
            break;
          nb = SUB;
          if (!done) {
            const double lo2 = ((yy - thr_g) - B.c4) * 0.5;
#pragma unroll
            for (int r = 0; r < SUB; r++) {
              const double* x = lds64 + (s0 + r) * ST;
              double e = x[EMAX];
#pragma unroll
              for (int j = 0; j < EMAX; j++)
                e = __builtin_fma(x[j], y[j], e);
              mask = shift_in(mask, e > lo2);
            }
          }
        }
        if (k == KCAP)
          walk_rows_pos<KCAP, EMAX, true>(mask, lds64 + s0 * ST, lpos + s0, y, k, la, li, kth, thr_g, zeros, near);
        else
          walk_rows_pos<KCAP, EMAX, false>(mask, lds64 + s0 * ST, lpos + s0, y, k, la, li, kth, thr_g, zeros, near);
        s0 += nb;
      }
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
      finish_target<KCAP>(B, tk, t, np, lrows, nl, zeros, k, ls, li, tgt, eth);
    } while (0);
    __syncthreads();
    item = item_sh;
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

// The CPU path's selection for one (task, target) on a wavefront: lane l takes library positions l,
// l + 64, ... and keeps its own k nearest by (sqrt distance, position) as they arrive; the wavefront
// then merges the 64 lists, smallest (sqrt distance, position) first, into the task's list (it holds
// the k nearest of the union), and lane 0 finishes the target.
template <int KCAP>
__device__ void exact_target_wave(const BatchParams& B, int tk, int t, int lane)
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
  // four of the lane's positions at a time, so that their loads are in flight together
  constexpr int RW = 4;
  for (int c0 = lane; c0 < nl; c0 += 64 * RW) {
    const double* xr[RW];
    double acc[RW];
#pragma unroll
    for (int u = 0; u < RW; u++) {
      const int c = c0 + 64 * u;
      xr[u] = B.L + (size_t)lrows[c < nl ? c : c0] * Ea;
      acc[u] = 0.0;
    }
    for (int j = 0; j < Ea; j++) {
      const double yj = yp[j];
#pragma unroll
      for (int u = 0; u < RW; u++) {
        const double d = xr[u][j] - yj;
        acc[u] += d * d;
      }
    }
#pragma unroll
    for (int u = 0; u < RW; u++) {
      const int c = c0 + 64 * u;
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
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    zeros += __shfl_xor(zeros, off);
  double ms[KCAP];
  int mi[KCAP];
#pragma unroll
  for (int r = 0; r < KCAP; r++) {
    // the smallest (sqrt distance, position) among the lanes' heads; positions are distinct
    double bs = ls[0];
    int bi = li[0], bl = lane;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double os = __shfl_xor(bs, off);
      const int oi = __shfl_xor(bi, off), ol = __shfl_xor(bl, off);
      const bool take = os < bs || (os == bs && (oi < bi || (oi == bi && ol < bl)));
      bs = take ? os : bs;
      bi = take ? oi : bi;
      bl = take ? ol : bl;
    }
    ms[r] = bs;
    mi[r] = bi;
    if (lane == bl) {
#pragma unroll
      for (int j = 0; j + 1 < KCAP; j++) {
        ls[j] = ls[j + 1];
        li[j] = li[j + 1];
      }
      ls[KCAP - 1] = INF_D;
      li[KCAP - 1] = INT_MAX;
    }
  }
  if (lane == 0)
    finish_target<KCAP>(B, tk, t, np, lrows, nl, zeros, k, ms, mi);
}

// The (task, target) pairs handed over (at most redo_cap of them), one per wavefront; or, when more
// pairs were handed over than the list holds, every target of every item of the batch's work lists.
template <int KCAP>
__global__ __launch_bounds__(256) void batch_redo(BatchParams B, RedoLists L)
{
  if (*B.redo_count > B.redo_cap) {
    for (int q = 0; q < 3; q++)
      for (int item = blockIdx.x; item < L.counts[q]; item += gridDim.x) {
        const int2 it = L.items[q][item];
        const int t = it.y + threadIdx.x;
        if (t < B.pred_count[it.x])
          exact_target<KCAP>(B, it.x, t);
      }
  } else {
    const int n = min(*B.redo_count, B.redo_cap);
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < n; r += waves) {
      const int2 e = B.redo[r];
      exact_target_wave<KCAP>(B, e.x, e.y, lane);
    }
  }
}

#define FEDM_BATCH(KC, EM)                                                                                       \
  template __global__ void batch_simplex<KC, EM, batch_waves(KC, EM)>(BatchParams);                              \
  template __global__ void batch_simplex<KC, EM, batch_waves_alt(KC, EM)>(BatchParams);
#define FEDM_BATCH_E(KC)                                                                                         \
  FEDM_BATCH(KC, 2) FEDM_BATCH(KC, 3) FEDM_BATCH(KC, 4) FEDM_BATCH(KC, 8) FEDM_BATCH(KC, 16) FEDM_BATCH(KC, 32)  \
  template __global__ void batch_redo<KC>(BatchParams, RedoLists);
FEDM_BATCH_E(4)
FEDM_BATCH_E(8)
FEDM_BATCH_E(16)
FEDM_BATCH_E(32)
#define FEDM_PRUNED(KC)                                                                                          \
  template __global__ void batch_simplex_pruned<KC, 2>(BatchParams);                                           \
  template __global__ void batch_simplex_pruned<KC, 3>(BatchParams);                                           \
  template __global__ void batch_simplex_pruned<KC, 4>(BatchParams);                                           \
  template __global__ void batch_simplex_pruned<KC, 8>(BatchParams);
FEDM_PRUNED(4)
FEDM_PRUNED(8)
FEDM_PRUNED(16)
FEDM_PRUNED(32)

} // namespace fedm
