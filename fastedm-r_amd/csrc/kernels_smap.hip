// S-map predictions (src/edm.cpp:737-814): several (target, theta) systems per wavefront.
//
// A persistent grid drains groups of G = min(8, 64 / n) (target, theta) items; item i of a call is
// target t0 + i / T at theta i % T.  Per item the wavefront forms X^T X and X^T y in FP64, each entry
// summed neighbour by neighbour in neighbour order (as the CPU path), into the item's LDS slot.  The G
// systems are then solved together, one lane segment of n lanes per system: every system runs the
// same cyclic (p, q) sweep order, so p and q are uniform over the wavefront while each segment decides
// its own rotations.  The solve is the two-sided Jacobi SVD of Eigen 3.4's JacobiSVD, operation for
// operation as in jacobi_wave.h, with W, U and V of every system in LDS.
#include <algorithm>
#include <type_traits>

#include "kernels_common.h"
#include "fdlibm_exp.h"

namespace fedm {

constexpr int SM_CH = 32; // neighbours staged per Gram chunk

__host__ __device__ inline int smap_group(int n)
{
  return n > 64 ? 0 : (64 / n > 8 ? 8 : 64 / n);
}

// LDS doubles of smap_multi: W [G n^2] | U, V [G n^2 each] | b [G n] | (n < 3 only) sv, tmp, x [G n each].
// Before the solve starts, U's and V's slots hold the Gram staging Xc [SM_CH n], yc [SM_CH], then acc
// [n (n+1)/2 + n], cd [SM_CH] and cidx [SM_CH] ints; after the sweeps, sv, tmp and x live in W's slot
// (n >= 3).  For n = 41 that is 40.7 KB, four wavefronts per CU.
__host__ __device__ inline size_t smap_uv_doubles(int n, int G)
{
  const size_t nacc = (size_t)n * (n + 1) / 2 + n;
  return std::max((size_t)2 * G * n * n, (size_t)SM_CH * n + SM_CH + nacc + SM_CH + SM_CH / 2);
}
__host__ __device__ inline size_t smap_lds_doubles(int n, int G)
{
  return (size_t)G * n * n + smap_uv_doubles(n, G) + (size_t)G * n + (n >= 3 ? 0 : 3 * (size_t)G * n);
}

// Segmented Jacobi SVD solve.  Lane g n + j (g < G) owns row/column j of system g.
__device__ void jacobi_svd_solve_multi(int n, int G, double* W, double* U, double* V, const double* b, double* sv,
                                       double* tmp, double* x, bool skip, int lane)
{
  const int g = lane / n, j = lane - (lane / n) * n;
  const bool act = g < G;
  const int gg = act ? g : 0;
  double* Wg = W + (size_t)gg * n * n;
  double* Ug = U + (size_t)gg * n * n;
  double* Vg = V + (size_t)gg * n * n;
  const double* bg = b + gg * n;
  double* svg = sv + gg * n;
  double* tmpg = tmp + gg * n;
  double* xg = x + gg * n;

  // scale by the largest |w| of the system
  double cm = 0.0;
  if (act)
    for (int i = 0; i < n; i++)
      cm = fmax(cm, fabs(Wg[i * n + j]));
  double scale = 0.0;
  for (int i = 0; i < n; i++)
    scale = fmax(scale, __shfl(cm, gg * n + i));
  const bool bad = skip || !act || !isfinite(scale);
  if (scale == 0.0)
    scale = 1.0;
  if (!bad) {
    for (int i = 0; i < n; i++)
      Wg[i * n + j] = Wg[i * n + j] / scale;
    for (int i = 0; i < n; i++) {
      Ug[j * n + i] = (i == j) ? 1.0 : 0.0;
      Vg[j * n + i] = (i == j) ? 1.0 : 0.0;
    }
  }
  __syncthreads();
  const double precision = 2.0 * DBL_EPSILON;
  double maxDiag = 0.0;
  if (!bad)
    for (int i = 0; i < n; i++)
      maxDiag = fmax(maxDiag, fabs(Wg[i * n + i]));

  bool finished = bad;
  int sweeps = 0;
  while (__ballot(!finished) != 0ull && sweeps < 1000) {
    sweeps++;
    const bool run = !finished;
    if (run)
      finished = true;
    for (int p = 1; p < n; p++) {
      for (int q = 0; q < p; q++) {
        bool rot = false;
        double wpq = 0.0, wqp = 0.0;
        if (run) {
          const double thr = fmax(DBL_MIN, precision * maxDiag);
          wpq = Wg[p * n + q];
          wqp = Wg[q * n + p];
          rot = fabs(wpq) > thr || fabs(wqp) > thr;
        }
        if (!__ballot(rot))
          continue;
        Rot jl{ 1.0, 0.0 }, jr{ 1.0, 0.0 };
        double up = 0.0, uq = 0.0, vp = 0.0, vq = 0.0;
        if (rot) {
          finished = false;
          up = Ug[j * n + p];
          uq = Ug[j * n + q];
          vp = Vg[j * n + p];
          vq = Vg[j * n + q];
          real_2x2(Wg[p * n + p], wpq, wqp, Wg[q * n + q], &jl, &jr);
        }
        __syncthreads();
        // rows p, q of W and columns p, q of U by jl
        if (rot && !(jl.c == 1.0 && jl.s == 0.0)) {
          const double xv = Wg[p * n + j], yv = Wg[q * n + j];
          Wg[p * n + j] = jl.c * xv + jl.s * yv;
          Wg[q * n + j] = -jl.s * xv + jl.c * yv;
          Ug[j * n + p] = jl.c * up + jl.s * uq;
          Ug[j * n + q] = -jl.s * up + jl.c * uq;
        }
        __syncthreads();
        // columns p, q of W and V by jr (applyOnTheRight: rotation (c, -s))
        const double tc = jr.c, ts = -jr.s;
        if (rot && !(tc == 1.0 && ts == 0.0)) {
          const double xv = Wg[j * n + p], yv = Wg[j * n + q];
          Wg[j * n + p] = tc * xv + ts * yv;
          Wg[j * n + q] = -ts * xv + tc * yv;
          Vg[j * n + p] = tc * vp + ts * vq;
          Vg[j * n + q] = -ts * vp + tc * vq;
        }
        __syncthreads();
        if (rot)
          maxDiag = fmax(maxDiag, fmax(fabs(Wg[p * n + p]), fabs(Wg[q * n + q])));
      }
    }
  }
  // non-negative singular values; each lane reads its diagonal entry first, since sv may share W's slot
  const double dj = bad ? 0.0 : Wg[j * n + j];
  __syncthreads();
  for (int i = 0; i < n; i++) {
    const double a = __shfl(dj, gg * n + i);
    if (!bad && a < 0.0)
      Ug[j * n + i] = -Ug[j * n + i];
  }
  if (!bad)
    svg[j] = fabs(dj) * scale;
  __syncthreads();
  // decreasing order, first maximum wins
  int nonzero = n;
  bool stop = bad;
  for (int i = 0; i < n; i++) {
    int pos = i;
    if (!stop) {
      double mx = svg[i];
      for (int c = i + 1; c < n; c++)
        if (svg[c] > mx) {
          mx = svg[c];
          pos = c;
        }
      if (mx == 0.0) {
        nonzero = i;
        stop = true;
      }
    }
    __syncthreads();
    if (!stop && pos != i) {
      if (j == 0) {
        const double a = svg[i];
        svg[i] = svg[pos];
        svg[pos] = a;
      }
      double a = Ug[j * n + i];
      Ug[j * n + i] = Ug[j * n + pos];
      Ug[j * n + pos] = a;
      a = Vg[j * n + i];
      Vg[j * n + i] = Vg[j * n + pos];
      Vg[j * n + pos] = a;
    }
    __syncthreads();
  }
  int rank = 0;
  if (!bad) {
    const double thr = fmax(svg[0] * ((double)n * DBL_EPSILON), DBL_MIN);
    int r = nonzero - 1;
    while (r >= 0 && svg[r] < thr)
      r--;
    rank = r + 1;
  }
  if (!bad && j < rank) {
    double s = 0.0;
    for (int c = 0; c < n; c++)
      s += Ug[c * n + j] * bg[c];
    tmpg[j] = (1.0 / svg[j]) * s;
  }
  __syncthreads();
  if (act) {
    double s = 0.0;
    for (int i = 0; i < rank; i++)
      s += Vg[j * n + i] * tmpg[i];
    xg[j] = s;
  }
  __syncthreads();
}

// A barrier over the group's buffers: the block's, or only the wavefront's when the block holds several
// wavefronts with buffers of their own (LDS operations of one wavefront complete in order).
template <bool WAVE>
__device__ __forceinline__ void group_sync()
{
  if (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  } else {
    __syncthreads();
  }
}

// Forms the systems of the group of G (target, theta) items that starts at item i0: X^T X into the slot
// W_g (both triangles) and X^T y into b_g, each entry summed neighbour by neighbour in neighbour order.
// Returns bit g set when item i0 + g has a system.  Weighted neighbour rows are staged sch at a time in
// stage, or, when stage is null, in the slot W_g of the system being formed (sch (n + 1) <= n^2).  With
// WAVE_SYNC the group's buffers belong to one wavefront of a larger block, which syncs on its own.
template <bool WAVE_SYNC = false>
__device__ unsigned form_group(const SmapParams& S, const double* __restrict__ D, int t0, int nt, int i0, int n,
                               int G, double* W, double* bvec, double* stage, int sch, double* acc, double* cd,
                               int* cidx, int lane)
{
  const int Ea = n - 1, T = S.T;
  const int ntri = n * (n + 1) / 2, nacc = ntri + n;
  const int items = nt * T;
  unsigned valid = 0; // bit g: item i0 + g has a system
  for (int g = 0; g < G; g++) {
    const int item = i0 + g;
    if (item >= items)
      break;
    const int tl = item / T, th = item - (item / T) * T;
    const int t = t0 + tl;
    if (t >= S.NP)
      continue;
    const int kk = S.nn_k[t];
    if (kk == 0)
      continue;
    valid |= 1u << g;
    if (kk < 0 && S.gram) {
      // formed by gram_dense
      const double* src = S.gram + (size_t)item * (n * n + n);
      double* Wg = W + (size_t)g * n * n;
      for (int e = lane; e < n * n; e += 64)
        Wg[e] = src[e];
      for (int e = lane; e < n; e += 64)
        bvec[g * n + e] = src[n * n + e];
      group_sync<WAVE_SYNC>();
      continue;
    }
    const bool dense = kk < 0;
    const int k = dense ? S.nn_cnt[t] : kk;
    const double* drow = dense ? D + (size_t)tl * S.NL : nullptr;
    const int* li = S.nn_i + (size_t)(t - S.nn_t0) * S.nn_stride;
    const double* lsd = S.nn_s + (size_t)(t - S.nn_t0) * S.nn_stride;
    double* Xc = stage ? stage : W + (size_t)g * n * n;
    double* yc = Xc + (size_t)sch * n;
    // mean of the selected distances, summed in neighbour order (Eigen mean = sum / size)
    double sum = 0.0;
    if (lane == 0) {
      if (dense) {
        for (int i = 0; i < S.NL; i++)
          if (drow[i] >= 0.0)
            sum += drow[i];
      } else {
        for (int j = 0; j < k; j++)
          sum += lsd[j];
      }
    }
    sum = __shfl(sum, 0);
    const double mean = sum / (double)k;
    const double theta = S.thetas[th];
    for (int e = lane; e < nacc; e += 64)
      acc[e] = 0.0;
    int pos = 0, done = 0;
    while (done < k) {
      int cnum = 0;
      if (dense) {
        // next run of valid entries of the distance row, in library order
        while (cnum == 0 && pos < S.NL) {
          const int i = pos + lane;
          const bool v = lane < sch && i < S.NL && drow[i] >= 0.0;
          const unsigned long long bb = __ballot(v);
          const int before = __popcll(bb & ((1ull << lane) - 1ull));
          if (v) {
            cidx[before] = i;
            cd[before] = drow[i];
          }
          cnum = __popcll(bb);
          pos += sch;
        }
      } else {
        cnum = min(sch, k - done);
        if (lane < cnum) {
          cidx[lane] = li[done + lane];
          cd[lane] = lsd[done + lane];
        }
      }
      if (cnum == 0)
        break;
      group_sync<WAVE_SYNC>();
      if (lane < cnum) {
        const int ci = cidx[lane];
        const double w = fdlibm_exp(-theta * (cd[lane] / mean));
        const double* row = S.M + (size_t)ci * Ea;
        Xc[lane * n] = w;
        for (int j = 1; j < n; j++)
          Xc[lane * n + j] = w * row[j - 1];
        yc[lane] = w * S.targets[ci];
      }
      group_sync<WAVE_SYNC>();
      for (int e = lane; e < nacc; e += 64) {
        double a = acc[e];
        if (e < ntri) {
          int p = 0, rem = e;
          while (rem >= n - p) {
            rem -= n - p;
            p++;
          }
          const int q = p + rem;
          for (int c = 0; c < cnum; c++)
            a += Xc[c * n + p] * Xc[c * n + q];
        } else {
          const int p = e - ntri;
          for (int c = 0; c < cnum; c++)
            a += Xc[c * n + p] * yc[c];
        }
        acc[e] = a;
      }
      group_sync<WAVE_SYNC>();
      done += cnum;
    }
    double* Wg = W + (size_t)g * n * n;
    for (int e = lane; e < nacc; e += 64) {
      if (e < ntri) {
        int p = 0, rem = e;
        while (rem >= n - p) {
          rem -= n - p;
          p++;
        }
        const int q = p + rem;
        Wg[p * n + q] = acc[e];
        Wg[q * n + p] = acc[e];
      } else {
        bvec[g * n + e - ntri] = acc[e];
      }
    }
    group_sync<WAVE_SYNC>();
  }
  return valid;
}

// Prediction (and, at the last theta, the coefficients) of each system of the group.
__device__ void write_group(const SmapParams& S, int t0, int i0, int n, int G, unsigned valid, const double* xsol,
                            int lane)
{
  const int Ea = n - 1, T = S.T;
  const int gl = lane / n;
  const bool mine = gl < G && ((valid >> gl) & 1u);
  if (!mine)
    return;
  const int item = i0 + gl;
  const int tl = item / T, th = item - (item / T) * T;
  const int t = t0 + tl;
  const int j = lane - gl * n;
  const double* xg = xsol + gl * n;
  if (j == 0) {
    double r = xg[0];
    const double* yp = S.Mp + (size_t)t * Ea;
    for (int c = 0; c < Ea; c++)
      if (yp[c] != MISSING_D)
        r += yp[c] * xg[c + 1];
    S.pred[(size_t)th * S.NP + t] = r;
    S.rc[(size_t)th * S.NP + t] = RC_SUCCESS;
  }
  if (S.coeffs && th == T - 1) {
    const double c = xg[j];
    S.coeffs[(size_t)t * n + j] = fabs(c) < 1.0e-11 ? MISSING_D : c;
  }
}

// Persistent grid over groups of G (target, theta) items of targets [t0, t0 + nt); U and V in LDS.
__global__ __launch_bounds__(64) void smap_multi(SmapParams S, const double* __restrict__ D, int t0, int nt)
{
  extern __shared__ double smem[];
  const int n = S.Ea + 1;
  const int G = smap_group(n);
  const int nacc = n * (n + 1) / 2 + n;
  const int lane = threadIdx.x;
  double* W = smem;
  double* U = W + (size_t)G * n * n;
  double* V = U + (size_t)G * n * n;
  double* bvec = U + smap_uv_doubles(n, G);
  double* sv = n >= 3 ? W : bvec + G * n;
  double* tmp = sv + G * n;
  double* xsol = tmp + G * n;
  double* acc = U + SM_CH * n + SM_CH; // after the staging
  double* cd = acc + nacc;
  int* cidx = (int*)(cd + SM_CH);

  const int items = nt * S.T;
  for (int i0 = blockIdx.x * G; i0 < items; i0 += gridDim.x * G) {
    // the Gram staging lives in U's slot until the solve starts
    const unsigned valid = form_group(S, D, t0, nt, i0, n, G, W, bvec, U, SM_CH, acc, cd, cidx, lane);
    const int gl = lane / n;
    const bool mine = gl < G && ((valid >> gl) & 1u);
    jacobi_svd_solve_multi(n, G, W, U, V, bvec, sv, tmp, xsol, !mine, lane);
    write_group(S, t0, i0, n, G, valid, xsol, lane);
    __syncthreads();
  }
}

// Element i (wave-uniform) of a row held in registers as NM + NT doubles.
template <int NM, int NT, class VM, class VT>
__device__ __forceinline__ double row_get(const VM& m, const VT& t, int i)
{
  if constexpr (NT == 0) {
    return m[i & (NM - 1)];
  } else {
    const double a = m[i & (NM - 1)], b = t[(i - NM) & (NT - 1)];
    return i < NM ? a : b;
  }
}

// Writes are unconditional (a select per element) so that the indexed register write never sits
// under a branch.
template <int NM, int NT, class VM, class VT>
__device__ __forceinline__ void row_set(VM& m, VT& t, int i, double v)
{
  if constexpr (NT == 0) {
    m[i & (NM - 1)] = v;
  } else {
    const int im = i & (NM - 1), it = (i - NM) & (NT - 1);
    const double om = m[im], ot = t[it];
    m[im] = i < NM ? v : om;
    t[it] = i < NM ? ot : v;
  }
}

// The solve of jacobi_svd_solve_multi with U and V in registers: lane g n + j keeps row j of U_g and
// of V_g as NM + NT doubles (n <= NM + NT).  Every rotation touches columns p and q of U and V, and p
// and q are wave-uniform, so each row element is reached by an indexed register move.  Element p is
// held in a plain register over the q loop.  The sort of the singular values keeps the column
// permutation in perm instead of moving columns; at the end the rows of U, then of V, go to W's slot
// for the two products of the solve.  Every floating-point operation and its order are those of
// jacobi_svd_solve_multi.
template <int NM, int NT>
__device__ void jacobi_svd_solve_rows(int n, int G, double* W, const double* b, double* sv, double* tmp,
                                      double* x, int* perm, bool skip, int lane)
{
  typedef double VM __attribute__((ext_vector_type(NM)));
  typedef double VT __attribute__((ext_vector_type(NT > 0 ? NT : 1)));
  constexpr int TM = NT > 0 ? NT - 1 : 0;
  const int g = lane / n, j = lane - (lane / n) * n;
  const bool act = g < G;
  const int gg = act ? g : 0;
  double* Wg = W + (size_t)gg * n * n;
  const double* bg = b + gg * n;
  double* svg = sv + gg * n;
  double* tmpg = tmp + gg * n;
  double* xg = x + gg * n;
  int* pg = perm + gg * n;

  // scale by the largest |w| of the system
  double cm = 0.0;
  if (act)
    for (int i = 0; i < n; i++)
      cm = fmax(cm, fabs(Wg[i * n + j]));
  double scale = 0.0;
  for (int i = 0; i < n; i++)
    scale = fmax(scale, __shfl(cm, gg * n + i));
  const bool bad = skip || !act || !isfinite(scale);
  if (scale == 0.0)
    scale = 1.0;
  if (!bad)
    for (int i = 0; i < n; i++)
      Wg[i * n + j] = Wg[i * n + j] / scale;
  // rows of U and V start as the identity; the opaque copy of j keeps the compiler from hoisting
  // these constants out of the item loop, where they would hold registers through the Gram phase
  int jo = j;
  asm volatile("" : "+v"(jo));
  VM um, vm;
  VT ut, vt;
#pragma unroll
  for (int i = 0; i < NM; i++) {
    um[i] = i == jo ? 1.0 : 0.0;
    vm[i] = um[i];
  }
#pragma unroll
  for (int i = 0; i < TM + 1; i++) {
    ut[i] = NM + i == jo ? 1.0 : 0.0;
    vt[i] = ut[i];
  }
  __syncthreads();
  const double precision = 2.0 * DBL_EPSILON;
  double maxDiag = 0.0;
  if (!bad)
    for (int i = 0; i < n; i++)
      maxDiag = fmax(maxDiag, fabs(Wg[i * n + i]));

  // The sweeps and their p loop run as one loop, so that the rows have a single loop-carried home in
  // registers: sweep s + 1 starts (p back to 1) while some system of the wavefront still rotated in
  // sweep s, at most 1000 sweeps.
  bool finished = bad;
  int sweeps = 0;
  bool go = __ballot(!finished) != 0ull;
  bool run = false;
  if (go) {
    sweeps = 1;
    run = !finished;
    if (run)
      finished = true;
  }
  int p = 1;
  while (go) {
    {
      double up = row_get<NM, NT>(um, ut, p), vp = row_get<NM, NT>(vm, vt, p);
      auto step = [&](int q, auto tail) {
        constexpr bool TAIL = decltype(tail)::value;
        bool rot = false;
        double wpq = 0.0, wqp = 0.0;
        if (run) {
          const double thr = fmax(DBL_MIN, precision * maxDiag);
          wpq = Wg[p * n + q];
          wqp = Wg[q * n + p];
          rot = fabs(wpq) > thr || fabs(wqp) > thr;
        }
        if (!__ballot(rot))
          return;
        Rot jl{ 1.0, 0.0 }, jr{ 1.0, 0.0 };
        if (rot) {
          finished = false;
          real_2x2(Wg[p * n + p], wpq, wqp, Wg[q * n + q], &jl, &jr);
        }
        __syncthreads();
        // rows p, q of W and columns p, q of U by jl
        const bool dl = rot && !(jl.c == 1.0 && jl.s == 0.0);
        if (dl) {
          const double xv = Wg[p * n + j], yv = Wg[q * n + j];
          Wg[p * n + j] = jl.c * xv + jl.s * yv;
          Wg[q * n + j] = -jl.s * xv + jl.c * yv;
        }
        {
          double uq;
          if constexpr (TAIL)
            uq = ut[(q - NM) & TM];
          else
            uq = um[q & (NM - 1)];
          const double np = jl.c * up + jl.s * uq, nq = -jl.s * up + jl.c * uq;
          up = dl ? np : up;
          if constexpr (TAIL)
            ut[(q - NM) & TM] = dl ? nq : uq;
          else
            um[q & (NM - 1)] = dl ? nq : uq;
        }
        __syncthreads();
        // columns p, q of W and V by jr (applyOnTheRight: rotation (c, -s))
        const double tc = jr.c, ts = -jr.s;
        const bool dr = rot && !(tc == 1.0 && ts == 0.0);
        if (dr) {
          const double xv = Wg[j * n + p], yv = Wg[j * n + q];
          Wg[j * n + p] = tc * xv + ts * yv;
          Wg[j * n + q] = -ts * xv + tc * yv;
        }
        {
          double vq;
          if constexpr (TAIL)
            vq = vt[(q - NM) & TM];
          else
            vq = vm[q & (NM - 1)];
          const double np = tc * vp + ts * vq, nq = -ts * vp + tc * vq;
          vp = dr ? np : vp;
          if constexpr (TAIL)
            vt[(q - NM) & TM] = dr ? nq : vq;
          else
            vm[q & (NM - 1)] = dr ? nq : vq;
        }
        __syncthreads();
        if (rot)
          maxDiag = fmax(maxDiag, fmax(fabs(Wg[p * n + p]), fabs(Wg[q * n + q])));
      };
      const int qm = p < NM ? p : NM;
      for (int q = 0; q < qm; q++)
        step(q, std::false_type{});
      if constexpr (NT > 0)
        for (int q = NM; q < p; q++)
          step(q, std::true_type{});
      row_set<NM, NT>(um, ut, p, up);
      row_set<NM, NT>(vm, vt, p, vp);
    }
    if (++p == n) {
      p = 1;
      go = __ballot(!finished) != 0ull && sweeps < 1000;
      if (go) {
        sweeps++;
        run = !finished;
        if (run)
          finished = true;
      }
    }
  }
  // non-negative singular values: bit i of flip negates column i of U (applied as the rows go out)
  unsigned flip = 0;
  for (int i = 0; i < n; i++) {
    const double a = bad ? 0.0 : Wg[i * n + i];
    if (!bad && a < 0.0)
      flip |= 1u << i;
    if (!bad && j == 0)
      svg[i] = fabs(a) * scale;
  }
  if (act)
    pg[j] = j;
  __syncthreads();
  // decreasing order, first maximum wins; perm[i] is the column of U and V now at position i
  int nonzero = n;
  bool stop = bad;
  for (int i = 0; i < n; i++) {
    int pos = i;
    if (!stop) {
      double mx = svg[i];
      for (int c = i + 1; c < n; c++)
        if (svg[c] > mx) {
          mx = svg[c];
          pos = c;
        }
      if (mx == 0.0) {
        nonzero = i;
        stop = true;
      }
    }
    __syncthreads();
    if (!stop && pos != i && j == 0) {
      const double a = svg[i];
      svg[i] = svg[pos];
      svg[pos] = a;
      const int c = pg[i];
      pg[i] = pg[pos];
      pg[pos] = c;
    }
    __syncthreads();
  }
  int rank = 0;
  if (!bad) {
    const double thr = fmax(svg[0] * ((double)n * DBL_EPSILON), DBL_MIN);
    int r = nonzero - 1;
    while (r >= 0 && svg[r] < thr)
      r--;
    rank = r + 1;
  }
  // rows of U into W's slot
  if (act) {
#pragma unroll
    for (int k = 0; k < NM; k++)
      if (k < n)
        Wg[j * n + k] = ((flip >> k) & 1u) ? -um[k] : um[k];
#pragma unroll
    for (int k = 0; k < NT; k++)
      if (NM + k < n)
        Wg[j * n + NM + k] = ((flip >> (NM + k)) & 1u) ? -ut[k] : ut[k];
  }
  __syncthreads();
  if (!bad && j < rank) {
    const int cj = pg[j];
    double s = 0.0;
    for (int c = 0; c < n; c++)
      s += Wg[c * n + cj] * bg[c];
    tmpg[j] = (1.0 / svg[j]) * s;
  }
  __syncthreads();
  // rows of V into W's slot
  if (act) {
#pragma unroll
    for (int k = 0; k < NM; k++)
      if (k < n)
        Wg[j * n + k] = vm[k];
#pragma unroll
    for (int k = 0; k < NT; k++)
      if (NM + k < n)
        Wg[j * n + NM + k] = vt[k];
  }
  __syncthreads();
  if (act) {
    double s = 0.0;
    for (int i = 0; i < rank; i++)
      s += Wg[j * n + pg[i]] * tmpg[i];
    xg[j] = s;
  }
  __syncthreads();
}

// Gram staging: in the slot of the system being formed when a chunk of at least 8 rows fits there.
__host__ __device__ inline int smap_rows_stage(int n)
{
  const int in_slot = n * n / (n + 1);
  return in_slot >= 8 ? (in_slot < SM_CH ? in_slot : SM_CH) : 0;
}

// LDS doubles of smap_rows: W [G n^2] | b, sv, tmp, x [G n each] | acc [n (n+1)/2 + n] | cd [SM_CH] |
// staging [SM_CH (n + 1)] when it does not fit a slot of W | cidx [SM_CH] and perm [G n] ints.
__host__ __device__ inline size_t smap_rows_lds_doubles(int n, int G)
{
  const size_t nacc = (size_t)n * (n + 1) / 2 + n;
  const size_t stage = smap_rows_stage(n) > 0 ? 0 : (size_t)SM_CH * (n + 1);
  return (size_t)G * n * n + 4 * (size_t)G * n + nacc + SM_CH + stage + (SM_CH + (size_t)G * n + 1) / 2;
}

// Persistent grid over groups of G (target, theta) items of targets [t0, t0 + nt); U and V rows in
// registers (n <= NM + NT), so a wavefront needs about a third of smap_multi's LDS.
template <int NM, int NT>
__global__ __launch_bounds__(64) void smap_rows(SmapParams S, const double* __restrict__ D, int t0, int nt)
{
  extern __shared__ double smem[];
  const int n = S.Ea + 1;
  const int G = smap_group(n);
  const int nacc = n * (n + 1) / 2 + n;
  const int lane = threadIdx.x;
  const int sin = smap_rows_stage(n);
  double* W = smem;
  double* bvec = W + (size_t)G * n * n;
  double* sv = bvec + G * n;
  double* tmp = sv + G * n;
  double* xsol = tmp + G * n;
  double* acc = xsol + G * n;
  double* cd = acc + nacc;
  double* stage = cd + SM_CH;
  int* cidx = (int*)(stage + (sin > 0 ? 0 : SM_CH * (n + 1)));
  int* perm = cidx + SM_CH;

  const int items = nt * S.T;
  for (int i0 = blockIdx.x * G; i0 < items; i0 += gridDim.x * G) {
    const unsigned valid =
        form_group(S, D, t0, nt, i0, n, G, W, bvec, sin > 0 ? nullptr : stage, sin > 0 ? sin : SM_CH, acc, cd,
                   cidx, lane);
    const int gl = lane / n;
    const bool mine = gl < G && ((valid >> gl) & 1u);
    jacobi_svd_solve_rows<NM, NT>(n, G, W, bvec, sv, tmp, xsol, perm, !mine, lane);
    write_group(S, t0, i0, n, G, valid, xsol, lane);
    __syncthreads();
  }
}

// Systems per wavefront of smap_pack: two row slots per lane (at most 32 systems, the width of the
// group's validity mask).
__host__ __device__ inline int smap_pack_group(int n)
{
  return 128 / n > 32 ? 32 : 128 / n;
}

// LDS doubles of smap_pack: smap_rows' buffers for G systems, then the step rotations [G x 4], the
// scales [G], and the ranks and sign masks [G ints each].
__host__ __device__ inline size_t smap_pack_lds_doubles(int n, int G)
{
  return smap_rows_lds_doubles(n, G) + 5 * (size_t)G + (size_t)G;
}

// The solve of jacobi_svd_solve_rows for up to 128 / n systems per wavefront.  Slot r (0, 1) of lane L is
// row j = (64 r + L) mod n of system g = (64 r + L) / n: the lane keeps row j of U_g and of V_g in
// registers and applies the rotations to column j (row update) and row j (column update) of W_g.  The
// scalar part of each rotation step, the threshold test and the 2x2 rotations, is made once per system:
// lane g < G owns system g, leaves the two rotations in LDS (the identity when the system does not
// rotate), and keeps the system's largest diagonal entry.  The scaling, the singular values, their sort
// and the rank are the owner lane's; the two products of the solve are the slot lanes'.  Every
// floating-point operation and its order are those of jacobi_svd_solve_multi.
template <int NM, int NT>
__device__ void jacobi_svd_solve_pack(int n, int G, double* W, const double* b, double* sv, double* tmp,
                                      double* x, int* perm, double* Rt, double* scl, int* rk, int* fl,
                                      unsigned live, int lane)
{
  typedef double VM __attribute__((ext_vector_type(NM)));
  typedef double VT __attribute__((ext_vector_type(NT > 0 ? NT : 1)));
  constexpr int TM = NT > 0 ? NT - 1 : 0;
  const int s1 = 64 + lane;
  const int g0 = lane / n, j0 = lane - (lane / n) * n;
  const int g1 = s1 / n, j1 = s1 - (s1 / n) * n;
  const bool act0 = g0 < G, act1 = g1 < G;
  const int gg0 = act0 ? g0 : 0, gg1 = act1 ? g1 : 0;
  double* W0 = W + (size_t)gg0 * n * n;
  double* W1 = W + (size_t)gg1 * n * n;

  // owner lanes: scale by the largest |w| of the system (the order of an exact maximum is immaterial)
  const bool A = lane < G;
  const int ga = A ? lane : 0;
  double* Wa = W + (size_t)ga * n * n;
  double* sva = sv + ga * n;
  int* pa = perm + ga * n;
  double scale = 0.0;
  if (A)
    for (int e = 0; e < n * n; e++)
      scale = fmax(scale, fabs(Wa[e]));
  const bool badA = !A || !((live >> lane) & 1u) || !isfinite(scale);
  if (scale == 0.0)
    scale = 1.0;
  if (A) {
    scl[lane] = scale;
    rk[lane] = badA ? -1 : 0;
  }
  group_sync<true>();
  const bool bad0 = !act0 || rk[gg0] < 0, bad1 = !act1 || rk[gg1] < 0;
  {
    const double sc0 = scl[gg0], sc1 = scl[gg1];
    if (!bad0)
      for (int i = 0; i < n; i++)
        W0[i * n + j0] = W0[i * n + j0] / sc0;
    if (!bad1)
      for (int i = 0; i < n; i++)
        W1[i * n + j1] = W1[i * n + j1] / sc1;
  }
  int jo0 = j0, jo1 = j1;
  asm volatile("" : "+v"(jo0), "+v"(jo1));
  VM um0, vm0, um1, vm1;
  VT ut0, vt0, ut1, vt1;
#pragma unroll
  for (int i = 0; i < NM; i++) {
    um0[i] = i == jo0 ? 1.0 : 0.0;
    vm0[i] = um0[i];
    um1[i] = i == jo1 ? 1.0 : 0.0;
    vm1[i] = um1[i];
  }
#pragma unroll
  for (int i = 0; i < TM + 1; i++) {
    ut0[i] = NM + i == jo0 ? 1.0 : 0.0;
    vt0[i] = ut0[i];
    ut1[i] = NM + i == jo1 ? 1.0 : 0.0;
    vt1[i] = ut1[i];
  }
  group_sync<true>();
  const double precision = 2.0 * DBL_EPSILON;
  double maxDiag = 0.0;
  if (!badA)
    for (int i = 0; i < n; i++)
      maxDiag = fmax(maxDiag, fabs(Wa[i * n + i]));

  // One rotation's update of a slot: rows p, q of W and columns p, q of U by (lc, ls), after a wavefront
  // sync columns p, q of W and V by (rc, -rs).
  auto left = [&](double* Ws, int j, bool act, double lc, double ls, int p, int q, double& up, double uq,
                  double& nq) {
    const bool dl = act && !(lc == 1.0 && ls == 0.0);
    if (dl) {
      const double xv = Ws[p * n + j], yv = Ws[q * n + j];
      Ws[p * n + j] = lc * xv + ls * yv;
      Ws[q * n + j] = -ls * xv + lc * yv;
    }
    const double np = lc * up + ls * uq, nn = -ls * up + lc * uq;
    up = dl ? np : up;
    nq = dl ? nn : uq;
  };
  auto right = [&](double* Ws, int j, bool act, double rc, double rs, int p, int q, double& vp, double vq,
                   double& nq) {
    const double tc = rc, ts = -rs;
    const bool dr = act && !(tc == 1.0 && ts == 0.0);
    if (dr) {
      const double xv = Ws[j * n + p], yv = Ws[j * n + q];
      Ws[j * n + p] = tc * xv + ts * yv;
      Ws[j * n + q] = -ts * xv + tc * yv;
    }
    const double np = tc * vp + ts * vq, nn = -ts * vp + tc * vq;
    vp = dr ? np : vp;
    nq = dr ? nn : vq;
  };

  // the sweeps and their p loop as one loop (one home for the register rows), at most 1000 sweeps
  bool finished = badA, run = false;
  int sweeps = 0;
  bool go = __ballot(!finished) != 0ull;
  if (go) {
    sweeps = 1;
    run = !finished;
    if (run)
      finished = true;
  }
  int p = 1;
  while (go) {
    {
      double up0 = row_get<NM, NT>(um0, ut0, p), vp0 = row_get<NM, NT>(vm0, vt0, p);
      double up1 = row_get<NM, NT>(um1, ut1, p), vp1 = row_get<NM, NT>(vm1, vt1, p);
      auto step = [&](int q, auto tail) {
        constexpr bool TAIL = decltype(tail)::value;
        bool rot = false;
        double wpq = 0.0, wqp = 0.0;
        if (run) {
          const double thr = fmax(DBL_MIN, precision * maxDiag);
          wpq = Wa[p * n + q];
          wqp = Wa[q * n + p];
          rot = fabs(wpq) > thr || fabs(wqp) > thr;
        }
        if (!__ballot(rot))
          return;
        if (A) {
          Rot jl{ 1.0, 0.0 }, jr{ 1.0, 0.0 };
          if (rot) {
            finished = false;
            real_2x2(Wa[p * n + p], wpq, wqp, Wa[q * n + q], &jl, &jr);
          }
          Rt[4 * lane + 0] = jl.c;
          Rt[4 * lane + 1] = jl.s;
          Rt[4 * lane + 2] = jr.c;
          Rt[4 * lane + 3] = jr.s;
        }
        group_sync<true>();
        const double lc0 = Rt[4 * gg0 + 0], ls0 = Rt[4 * gg0 + 1], rc0 = Rt[4 * gg0 + 2], rs0 = Rt[4 * gg0 + 3];
        const double lc1 = Rt[4 * gg1 + 0], ls1 = Rt[4 * gg1 + 1], rc1 = Rt[4 * gg1 + 2], rs1 = Rt[4 * gg1 + 3];
        if constexpr (TAIL) {
          double n0, n1;
          left(W0, j0, act0, lc0, ls0, p, q, up0, ut0[(q - NM) & TM], n0);
          ut0[(q - NM) & TM] = n0;
          left(W1, j1, act1, lc1, ls1, p, q, up1, ut1[(q - NM) & TM], n1);
          ut1[(q - NM) & TM] = n1;
        } else {
          double n0, n1;
          left(W0, j0, act0, lc0, ls0, p, q, up0, um0[q & (NM - 1)], n0);
          um0[q & (NM - 1)] = n0;
          left(W1, j1, act1, lc1, ls1, p, q, up1, um1[q & (NM - 1)], n1);
          um1[q & (NM - 1)] = n1;
        }
        group_sync<true>();
        if constexpr (TAIL) {
          double n0, n1;
          right(W0, j0, act0, rc0, rs0, p, q, vp0, vt0[(q - NM) & TM], n0);
          vt0[(q - NM) & TM] = n0;
          right(W1, j1, act1, rc1, rs1, p, q, vp1, vt1[(q - NM) & TM], n1);
          vt1[(q - NM) & TM] = n1;
        } else {
          double n0, n1;
          right(W0, j0, act0, rc0, rs0, p, q, vp0, vm0[q & (NM - 1)], n0);
          vm0[q & (NM - 1)] = n0;
          right(W1, j1, act1, rc1, rs1, p, q, vp1, vm1[q & (NM - 1)], n1);
          vm1[q & (NM - 1)] = n1;
        }
        group_sync<true>();
        if (rot)
          maxDiag = fmax(maxDiag, fmax(fabs(Wa[p * n + p]), fabs(Wa[q * n + q])));
      };
      const int qm = p < NM ? p : NM;
      for (int q = 0; q < qm; q++)
        step(q, std::false_type{});
      if constexpr (NT > 0)
        for (int q = NM; q < p; q++)
          step(q, std::true_type{});
      row_set<NM, NT>(um0, ut0, p, up0);
      row_set<NM, NT>(vm0, vt0, p, vp0);
      row_set<NM, NT>(um1, ut1, p, up1);
      row_set<NM, NT>(vm1, vt1, p, vp1);
    }
    if (++p == n) {
      p = 1;
      go = __ballot(!finished) != 0ull && sweeps < 1000;
      if (go) {
        sweeps++;
        run = !finished;
        if (run)
          finished = true;
      }
    }
  }
  // owner lanes: non-negative singular values (bit i of the sign mask negates column i of U), their
  // decreasing order (first maximum wins; perm[i] is the column now at position i) and the rank
  if (A) {
    unsigned flip = 0;
    for (int i = 0; i < n; i++) {
      const double a = badA ? 0.0 : Wa[i * n + i];
      if (!badA && a < 0.0)
        flip |= 1u << i;
      if (!badA)
        sva[i] = fabs(a) * scale;
      pa[i] = i;
    }
    int nonzero = n;
    if (!badA)
      for (int i = 0; i < n; i++) {
        int pos = i;
        double mx = sva[i];
        for (int c = i + 1; c < n; c++)
          if (sva[c] > mx) {
            mx = sva[c];
            pos = c;
          }
        if (mx == 0.0) {
          nonzero = i;
          break;
        }
        if (pos != i) {
          const double a = sva[i];
          sva[i] = sva[pos];
          sva[pos] = a;
          const int c = pa[i];
          pa[i] = pa[pos];
          pa[pos] = c;
        }
      }
    int rank = 0;
    if (!badA) {
      const double thr = fmax(sva[0] * ((double)n * DBL_EPSILON), DBL_MIN);
      int r = nonzero - 1;
      while (r >= 0 && sva[r] < thr)
        r--;
      rank = r + 1;
    }
    rk[lane] = rank;
    fl[lane] = (int)flip;
  }
  group_sync<true>();
  // rows of U (signs applied) into W's slot, then tmp = S^-1 (U^T b) over the rank
  const unsigned f0 = (unsigned)fl[gg0], f1 = (unsigned)fl[gg1];
  if (act0) {
#pragma unroll
    for (int k = 0; k < NM; k++)
      if (k < n)
        W0[j0 * n + k] = ((f0 >> k) & 1u) ? -um0[k] : um0[k];
#pragma unroll
    for (int k = 0; k < NT; k++)
      if (NM + k < n)
        W0[j0 * n + NM + k] = ((f0 >> (NM + k)) & 1u) ? -ut0[k] : ut0[k];
  }
  if (act1) {
#pragma unroll
    for (int k = 0; k < NM; k++)
      if (k < n)
        W1[j1 * n + k] = ((f1 >> k) & 1u) ? -um1[k] : um1[k];
#pragma unroll
    for (int k = 0; k < NT; k++)
      if (NM + k < n)
        W1[j1 * n + NM + k] = ((f1 >> (NM + k)) & 1u) ? -ut1[k] : ut1[k];
  }
  group_sync<true>();
  const int rank0 = act0 ? rk[gg0] : 0, rank1 = act1 ? rk[gg1] : 0;
  auto utb = [&](const double* Ws, int g, int j, int rank) {
    if (j < rank) {
      const int cj = perm[g * n + j];
      const double* bg = b + g * n;
      double s = 0.0;
      for (int c = 0; c < n; c++)
        s += Ws[c * n + cj] * bg[c];
      tmp[g * n + j] = (1.0 / sv[g * n + j]) * s;
    }
  };
  utb(W0, gg0, j0, rank0);
  utb(W1, gg1, j1, rank1);
  group_sync<true>();
  // rows of V into W's slot, then x = V tmp over the rank
  if (act0) {
#pragma unroll
    for (int k = 0; k < NM; k++)
      if (k < n)
        W0[j0 * n + k] = vm0[k];
#pragma unroll
    for (int k = 0; k < NT; k++)
      if (NM + k < n)
        W0[j0 * n + NM + k] = vt0[k];
  }
  if (act1) {
#pragma unroll
    for (int k = 0; k < NM; k++)
      if (k < n)
        W1[j1 * n + k] = vm1[k];
#pragma unroll
    for (int k = 0; k < NT; k++)
      if (NM + k < n)
        W1[j1 * n + NM + k] = vt1[k];
  }
  group_sync<true>();
  auto vx = [&](const double* Ws, int g, int j, int rank) {
    double s = 0.0;
    for (int i = 0; i < rank; i++)
      s += Ws[j * n + perm[g * n + i]] * tmp[g * n + i];
    x[g * n + j] = s;
  };
  if (act0)
    vx(W0, gg0, j0, rank0);
  if (act1)
    vx(W1, gg1, j1, rank1);
  group_sync<true>();
}

// Prediction (and, at the last theta, coefficient j) of system g of the group that starts at item i0.
__device__ void write_slot(const SmapParams& S, int t0, int i0, int n, int g, int j, const double* xsol)
{
  const int Ea = n - 1, T = S.T;
  const int item = i0 + g;
  const int tl = item / T, th = item - (item / T) * T;
  const int t = t0 + tl;
  const double* xg = xsol + g * n;
  if (j == 0) {
    double r = xg[0];
    const double* yp = S.Mp + (size_t)t * Ea;
    for (int c = 0; c < Ea; c++)
      if (yp[c] != MISSING_D)
        r += yp[c] * xg[c + 1];
    S.pred[(size_t)th * S.NP + t] = r;
    S.rc[(size_t)th * S.NP + t] = RC_SUCCESS;
  }
  if (S.coeffs && th == T - 1) {
    const double c = xg[j];
    S.coeffs[(size_t)t * n + j] = fabs(c) < 1.0e-11 ? MISSING_D : c;
  }
}

// Persistent grid over groups of G (target, theta) items (G <= smap_pack_group(n), chosen by the host),
// two row slots per lane.
template <int NM, int NT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void smap_pack(SmapParams S, const double* __restrict__ D, int t0, int nt, int G)
{
  extern __shared__ double smem[];
  const int n = S.Ea + 1;
  const int nacc = n * (n + 1) / 2 + n;
  const int lane = threadIdx.x;
  const int sin = smap_rows_stage(n);
  double* W = smem;
  double* bvec = W + (size_t)G * n * n;
  double* sv = bvec + G * n;
  double* tmp = sv + G * n;
  double* xsol = tmp + G * n;
  double* acc = xsol + G * n;
  double* cd = acc + nacc;
  double* stage = cd + SM_CH;
  int* cidx = (int*)(stage + (sin > 0 ? 0 : SM_CH * (n + 1)));
  int* perm = cidx + SM_CH;
  double* Rt = smem + smap_rows_lds_doubles(n, G);
  double* scl = Rt + 4 * G;
  int* rk = (int*)(scl + G);
  int* fl = rk + G;

  const int s1 = 64 + lane;
  const int g0 = lane / n, j0 = lane - (lane / n) * n;
  const int g1 = s1 / n, j1 = s1 - (s1 / n) * n;
  const int items = nt * S.T;
  for (int i0 = blockIdx.x * G; i0 < items; i0 += gridDim.x * G) {
    const unsigned valid =
        form_group(S, D, t0, nt, i0, n, G, W, bvec, sin > 0 ? nullptr : stage, sin > 0 ? sin : SM_CH, acc, cd,
                   cidx, lane);
    jacobi_svd_solve_pack<NM, NT>(n, G, W, bvec, sv, tmp, xsol, perm, Rt, scl, rk, fl, valid, lane);
    if (g0 < G && ((valid >> g0) & 1u))
      write_slot(S, t0, i0, n, g0, j0, xsol);
    if (g1 < G && ((valid >> g1) & 1u))
      write_slot(S, t0, i0, n, g1, j1, xsol);
    __syncthreads();
  }
}

// The rotations of one Jacobi step of one system of a block, made by the block's 2x2 lanes.
struct BlkRot
{
  double lc, ls, rc, rs;
  int rot, pad;
};

// jacobi_svd_solve_rows over a block of WB wavefronts, each with G systems in lane segments (U and V rows in
// registers, W in LDS), for up to 8 WB systems.  The 2x2 step of each rotation, a chain of IEEE divides
// and square roots that every lane of a segment would otherwise repeat, is made once per system: lane s
// of wavefront 0 takes system s of the block (its threshold test, its 2x2 rotations, its largest
// diagonal entry) and leaves the rotations in LDS; then every lane applies them to its column of W and
// its rows of U and V.  The block goes through the sweeps together (a system that has finished makes no
// rotations, as in the wavefront solve), so every system sees the operations of jacobi_svd_solve_multi in
// its order.  The scaling before and the sort and products after are the wavefront's own.
template <int NM, int NT, int WB>
__device__ void jacobi_svd_solve_blk(int n, int G, double* Wall, const double* b, double* sv, double* tmp,
                                     double* x, int* perm, bool skip, int lane, int w, BlkRot* R, int* sbad)
{
  typedef double VM __attribute__((ext_vector_type(NM)));
  typedef double VT __attribute__((ext_vector_type(NT > 0 ? NT : 1)));
  constexpr int TM = NT > 0 ? NT - 1 : 0;
  const int g = lane / n, j = lane - (lane / n) * n;
  const bool act = g < G;
  const int gg = act ? g : 0;
  double* Wg = Wall + (size_t)(w * G + gg) * n * n;
  const BlkRot* Rg = R + w * G + gg;
  const double* bg = b + gg * n;
  double* svg = sv + gg * n;
  double* tmpg = tmp + gg * n;
  double* xg = x + gg * n;
  int* pg = perm + gg * n;

  // scale by the largest |w| of the system
  double cm = 0.0;
  if (act)
    for (int i = 0; i < n; i++)
      cm = fmax(cm, fabs(Wg[i * n + j]));
  double scale = 0.0;
  for (int i = 0; i < n; i++)
    scale = fmax(scale, __shfl(cm, gg * n + i));
  const bool bad = skip || !act || !isfinite(scale);
  if (scale == 0.0)
    scale = 1.0;
  if (!bad)
    for (int i = 0; i < n; i++)
      Wg[i * n + j] = Wg[i * n + j] / scale;
  if (act && j == 0)
    sbad[w * G + g] = bad ? 1 : 0;
  int jo = j;
  asm volatile("" : "+v"(jo));
  VM um, vm;
  VT ut, vt;
#pragma unroll
  for (int i = 0; i < NM; i++) {
    um[i] = i == jo ? 1.0 : 0.0;
    vm[i] = um[i];
  }
#pragma unroll
  for (int i = 0; i < TM + 1; i++) {
    ut[i] = NM + i == jo ? 1.0 : 0.0;
    vt[i] = ut[i];
  }
  __syncthreads();
  // the 2x2 lanes: lane s of wavefront 0 owns system s of the block
  const bool A = w == 0 && lane < WB * G;
  double* Ws = Wall + (size_t)(A ? lane : 0) * n * n;
  const double precision = 2.0 * DBL_EPSILON;
  double maxDiag = 0.0;
  bool finished = true, run = false;
  if (A) {
    finished = sbad[lane] != 0;
    if (!finished)
      for (int i = 0; i < n; i++)
        maxDiag = fmax(maxDiag, fabs(Ws[i * n + i]));
  }
  // sweeps and their p loop as one loop (one home for the register rows), at most 1000 sweeps
  int sweeps = 0;
  bool go = __syncthreads_or(A && !finished) != 0;
  if (go) {
    sweeps = 1;
    run = !finished;
    if (run)
      finished = true;
  }
  int p = 1;
  while (go) {
    {
      double up = row_get<NM, NT>(um, ut, p), vp = row_get<NM, NT>(vm, vt, p);
      auto step = [&](int q, auto tail) {
        constexpr bool TAIL = decltype(tail)::value;
        int rot = 0;
        if (A) {
          double wpq = 0.0, wqp = 0.0;
          if (run) {
            const double thr = fmax(DBL_MIN, precision * maxDiag);
            wpq = Ws[p * n + q];
            wqp = Ws[q * n + p];
            rot = fabs(wpq) > thr || fabs(wqp) > thr;
          }
          Rot jl{ 1.0, 0.0 }, jr{ 1.0, 0.0 };
          if (rot) {
            finished = false;
            real_2x2(Ws[p * n + p], wpq, wqp, Ws[q * n + q], &jl, &jr);
          }
          BlkRot* r = R + lane;
          r->lc = jl.c;
          r->ls = jl.s;
          r->rc = jr.c;
          r->rs = jr.s;
          r->rot = rot;
        }
        if (!__syncthreads_or(rot))
          return;
        const bool me = act && Rg->rot != 0;
        const double lc = Rg->lc, ls = Rg->ls, rc = Rg->rc, rs = Rg->rs;
        // rows p, q of W and columns p, q of U by jl
        const bool dl = me && !(lc == 1.0 && ls == 0.0);
        if (dl) {
          const double xv = Wg[p * n + j], yv = Wg[q * n + j];
          Wg[p * n + j] = lc * xv + ls * yv;
          Wg[q * n + j] = -ls * xv + lc * yv;
        }
        {
          double uq;
          if constexpr (TAIL)
            uq = ut[(q - NM) & TM];
          else
            uq = um[q & (NM - 1)];
          const double np = lc * up + ls * uq, nq = -ls * up + lc * uq;
          up = dl ? np : up;
          if constexpr (TAIL)
            ut[(q - NM) & TM] = dl ? nq : uq;
          else
            um[q & (NM - 1)] = dl ? nq : uq;
        }
        __syncthreads();
        // columns p, q of W and V by jr (applyOnTheRight: rotation (c, -s))
        const double tc = rc, ts = -rs;
        const bool dr = me && !(tc == 1.0 && ts == 0.0);
        if (dr) {
          const double xv = Wg[j * n + p], yv = Wg[j * n + q];
          Wg[j * n + p] = tc * xv + ts * yv;
          Wg[j * n + q] = -ts * xv + tc * yv;
        }
        {
          double vq;
          if constexpr (TAIL)
            vq = vt[(q - NM) & TM];
          else
            vq = vm[q & (NM - 1)];
          const double np = tc * vp + ts * vq, nq = -ts * vp + tc * vq;
          vp = dr ? np : vp;
          if constexpr (TAIL)
            vt[(q - NM) & TM] = dr ? nq : vq;
          else
            vm[q & (NM - 1)] = dr ? nq : vq;
        }
        __syncthreads();
        if (A && rot)
          maxDiag = fmax(maxDiag, fmax(fabs(Ws[p * n + p]), fabs(Ws[q * n + q])));
      };
      const int qm = p < NM ? p : NM;
      for (int q = 0; q < qm; q++)
        step(q, std::false_type{});
      if constexpr (NT > 0)
        for (int q = NM; q < p; q++)
          step(q, std::true_type{});
      row_set<NM, NT>(um, ut, p, up);
      row_set<NM, NT>(vm, vt, p, vp);
    }
    if (++p == n) {
      p = 1;
      go = __syncthreads_or(A && !finished) != 0 && sweeps < 1000;
      if (go) {
        sweeps++;
        run = !finished;
        if (run)
          finished = true;
      }
    }
  }
  // from here on each wavefront works on its own systems
  unsigned flip = 0;
  for (int i = 0; i < n; i++) {
    const double a = bad ? 0.0 : Wg[i * n + i];
    if (!bad && a < 0.0)
      flip |= 1u << i;
    if (!bad && j == 0)
      svg[i] = fabs(a) * scale;
  }
  if (act)
    pg[j] = j;
  group_sync<true>();
  int nonzero = n;
  bool stop = bad;
  for (int i = 0; i < n; i++) {
    int pos = i;
    if (!stop) {
      double mx = svg[i];
      for (int c = i + 1; c < n; c++)
        if (svg[c] > mx) {
          mx = svg[c];
          pos = c;
        }
      if (mx == 0.0) {
        nonzero = i;
        stop = true;
      }
    }
    group_sync<true>();
    if (!stop && pos != i && j == 0) {
      const double a = svg[i];
      svg[i] = svg[pos];
      svg[pos] = a;
      const int c = pg[i];
      pg[i] = pg[pos];
      pg[pos] = c;
    }
    group_sync<true>();
  }
  int rank = 0;
  if (!bad) {
    const double thr = fmax(svg[0] * ((double)n * DBL_EPSILON), DBL_MIN);
    int r = nonzero - 1;
    while (r >= 0 && svg[r] < thr)
      r--;
    rank = r + 1;
  }
  if (act) {
#pragma unroll
    for (int k = 0; k < NM; k++)
      if (k < n)
        Wg[j * n + k] = ((flip >> k) & 1u) ? -um[k] : um[k];
#pragma unroll
    for (int k = 0; k < NT; k++)
      if (NM + k < n)
        Wg[j * n + NM + k] = ((flip >> (NM + k)) & 1u) ? -ut[k] : ut[k];
  }
  group_sync<true>();
  if (!bad && j < rank) {
    const int cj = pg[j];
    double s = 0.0;
    for (int c = 0; c < n; c++)
      s += Wg[c * n + cj] * bg[c];
    tmpg[j] = (1.0 / svg[j]) * s;
  }
  group_sync<true>();
  if (act) {
#pragma unroll
    for (int k = 0; k < NM; k++)
      if (k < n)
        Wg[j * n + k] = vm[k];
#pragma unroll
    for (int k = 0; k < NT; k++)
      if (NM + k < n)
        Wg[j * n + NM + k] = vt[k];
  }
  group_sync<true>();
  if (act) {
    double s = 0.0;
    for (int i = 0; i < rank; i++)
      s += Wg[j * n + pg[i]] * tmpg[i];
    xg[j] = s;
  }
  group_sync<true>();
}

// LDS of smap_blk, in doubles: the block's W [WB G n^2], the step rotations and flags, then per wavefront
// smap_rows' buffers without W.
constexpr int BLK_WAVES = 4;
__host__ __device__ inline size_t smap_blk_wave_doubles(int n, int G)
{
  return smap_rows_lds_doubles(n, G) - (size_t)G * n * n;
}
__host__ __device__ inline size_t smap_blk_lds_doubles(int n, int G, int WB)
{
  const size_t rot = ((size_t)WB * G * sizeof(BlkRot) + (size_t)WB * G * sizeof(int) + 7) / 8;
  return (size_t)WB * G * n * n + rot + (size_t)WB * smap_blk_wave_doubles(n, G);
}

// smap_rows with the block solve: WB wavefronts of G systems each; wavefront w forms and finishes items
// base + w G .. base + w G + G - 1 of each block group on its own, and the solve runs block-wide.
template <int NM, int NT, int WB>
__global__ __launch_bounds__(64 * WB) void smap_blk(SmapParams S, const double* __restrict__ D, int t0, int nt)
{
  extern __shared__ double smem[];
  const int n = S.Ea + 1;
  const int G = smap_group(n);
  const int nacc = n * (n + 1) / 2 + n;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int sin = smap_rows_stage(n);
  double* Wall = smem;
  BlkRot* R = (BlkRot*)(Wall + (size_t)WB * G * n * n);
  int* sbad = (int*)(R + WB * G);
  double* base_w = smem + (smap_blk_lds_doubles(n, G, WB) - (size_t)WB * smap_blk_wave_doubles(n, G)) +
                   (size_t)w * smap_blk_wave_doubles(n, G);
  double* bvec = base_w;
  double* sv = bvec + G * n;
  double* tmp = sv + G * n;
  double* xsol = tmp + G * n;
  double* acc = xsol + G * n;
  double* cd = acc + nacc;
  double* stage = cd + SM_CH;
  int* cidx = (int*)(stage + (sin > 0 ? 0 : SM_CH * (n + 1)));
  int* perm = cidx + SM_CH;
  double* W = Wall + (size_t)w * G * n * n;

  const int items = nt * S.T;
  for (int base = blockIdx.x * WB * G; base < items; base += gridDim.x * WB * G) {
    const int i0 = base + w * G;
    const unsigned valid = form_group<true>(S, D, t0, nt, i0, n, G, W, bvec, sin > 0 ? nullptr : stage,
                                            sin > 0 ? sin : SM_CH, acc, cd, cidx, lane);
    const int gl = lane / n;
    const bool mine = gl < G && ((valid >> gl) & 1u);
    jacobi_svd_solve_blk<NM, NT, WB>(n, G, Wall, bvec, sv, tmp, xsol, perm, !mine, lane, w, R, sbad);
    write_group(S, t0, i0, n, G, valid, xsol, lane);
    __syncthreads();
  }
}

// One system per wavefront (33 <= n <= 48): lane j keeps row j of U in registers as three 16-wide
// segments and row j of V in LDS with W.  The operations and their order are those of
// jacobi_svd_solve_multi, as in jacobi_svd_solve_rows.
__device__ void jacobi_svd_solve_hyb(int n, double* W, double* V, const double* b, double* sv, double* tmp,
                                     double* x, int* perm, bool skip, int lane)
{
  typedef double V16 __attribute__((ext_vector_type(16)));
  const int j = lane;
  const bool act = j < n;
  // scale by the largest |w| of the system
  double cm = 0.0;
  if (act)
    for (int i = 0; i < n; i++)
      cm = fmax(cm, fabs(W[i * n + j]));
  double scale = 0.0;
  for (int i = 0; i < n; i++)
    scale = fmax(scale, __shfl(cm, i));
  const bool bad = skip || !act || !isfinite(scale);
  if (scale == 0.0)
    scale = 1.0;
  if (!bad)
    for (int i = 0; i < n; i++)
      W[i * n + j] = W[i * n + j] / scale;
  // U rows in registers (identity); the opaque copy of j keeps the constants from being hoisted
  int jo = j;
  asm volatile("" : "+v"(jo));
  V16 u0, u1, u2;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    u0[i] = i == jo ? 1.0 : 0.0;
    u1[i] = 16 + i == jo ? 1.0 : 0.0;
    u2[i] = 32 + i == jo ? 1.0 : 0.0;
  }
  if (act)
    for (int i = 0; i < n; i++)
      V[j * n + i] = i == j ? 1.0 : 0.0;
  __syncthreads();
  auto uget = [&](int i) {
    const double a0 = u0[i & 15], a1 = u1[i & 15], a2 = u2[i & 15];
    return i < 16 ? a0 : (i < 32 ? a1 : a2);
  };
  auto uset = [&](int i, double v) {
    const int ii = i & 15;
    const double o0 = u0[ii], o1 = u1[ii], o2 = u2[ii];
    u0[ii] = i < 16 ? v : o0;
    u1[ii] = (i >= 16 && i < 32) ? v : o1;
    u2[ii] = i >= 32 ? v : o2;
  };
  const double precision = 2.0 * DBL_EPSILON;
  double maxDiag = 0.0;
  if (!bad)
    for (int i = 0; i < n; i++)
      maxDiag = fmax(maxDiag, fabs(W[i * n + i]));

  bool finished = bad;
  int sweeps = 0;
  bool go = __ballot(!finished) != 0ull;
  bool run = false;
  if (go) {
    sweeps = 1;
    run = !finished;
    if (run)
      finished = true;
  }
  int p = 1;
  while (go) {
    {
      double up = uget(p);
      auto step = [&](int q, auto seg) {
        constexpr int S = decltype(seg)::value;
        bool rot = false;
        double wpq = 0.0, wqp = 0.0;
        if (run) {
          const double thr = fmax(DBL_MIN, precision * maxDiag);
          wpq = W[p * n + q];
          wqp = W[q * n + p];
          rot = fabs(wpq) > thr || fabs(wqp) > thr;
        }
        if (!__ballot(rot))
          return;
        Rot jl{ 1.0, 0.0 }, jr{ 1.0, 0.0 };
        double vp = 0.0, vq = 0.0;
        if (rot) {
          finished = false;
          vp = V[j * n + p];
          vq = V[j * n + q];
          real_2x2(W[p * n + p], wpq, wqp, W[q * n + q], &jl, &jr);
        }
        __syncthreads();
        // rows p, q of W and columns p, q of U by jl
        const bool dl = rot && !(jl.c == 1.0 && jl.s == 0.0);
        if (dl) {
          const double xv = W[p * n + j], yv = W[q * n + j];
          W[p * n + j] = jl.c * xv + jl.s * yv;
          W[q * n + j] = -jl.s * xv + jl.c * yv;
        }
        {
          V16& us = S == 0 ? u0 : (S == 1 ? u1 : u2);
          const double uq = us[q & 15];
          const double np = jl.c * up + jl.s * uq, nq = -jl.s * up + jl.c * uq;
          up = dl ? np : up;
          us[q & 15] = dl ? nq : uq;
        }
        __syncthreads();
        // columns p, q of W and V by jr (applyOnTheRight: rotation (c, -s))
        const double tc = jr.c, ts = -jr.s;
        if (rot && !(tc == 1.0 && ts == 0.0)) {
          const double xv = W[j * n + p], yv = W[j * n + q];
          W[j * n + p] = tc * xv + ts * yv;
          W[j * n + q] = -ts * xv + tc * yv;
          V[j * n + p] = tc * vp + ts * vq;
          V[j * n + q] = -ts * vp + tc * vq;
        }
        __syncthreads();
        if (rot)
          maxDiag = fmax(maxDiag, fmax(fabs(W[p * n + p]), fabs(W[q * n + q])));
      };
      const int q1 = p < 16 ? p : 16, q2 = p < 32 ? p : 32;
      for (int q = 0; q < q1; q++)
        step(q, std::integral_constant<int, 0>{});
      for (int q = 16; q < q2; q++)
        step(q, std::integral_constant<int, 1>{});
      for (int q = 32; q < p; q++)
        step(q, std::integral_constant<int, 2>{});
      uset(p, up);
    }
    if (++p == n) {
      p = 1;
      go = __ballot(!finished) != 0ull && sweeps < 1000;
      if (go) {
        sweeps++;
        run = !finished;
        if (run)
          finished = true;
      }
    }
  }
  // non-negative singular values: bit i of flip negates column i of U (applied as the rows go out)
  unsigned long long flip = 0;
  for (int i = 0; i < n; i++) {
    const double a = bad ? 0.0 : W[i * n + i];
    if (!bad && a < 0.0)
      flip |= 1ull << i;
    if (!bad && j == 0)
      sv[i] = fabs(a) * scale;
  }
  if (act)
    perm[j] = j;
  __syncthreads();
  // decreasing order, first maximum wins; perm[i] is the column of U and V now at position i
  int nonzero = n;
  bool stop = bad;
  for (int i = 0; i < n; i++) {
    int pos = i;
    if (!stop) {
      double mx = sv[i];
      for (int c = i + 1; c < n; c++)
        if (sv[c] > mx) {
          mx = sv[c];
          pos = c;
        }
      if (mx == 0.0) {
        nonzero = i;
        stop = true;
      }
    }
    __syncthreads();
    if (!stop && pos != i && j == 0) {
      const double a = sv[i];
      sv[i] = sv[pos];
      sv[pos] = a;
      const int c = perm[i];
      perm[i] = perm[pos];
      perm[pos] = c;
    }
    __syncthreads();
  }
  int rank = 0;
  if (!bad) {
    const double thr = fmax(sv[0] * ((double)n * DBL_EPSILON), DBL_MIN);
    int r = nonzero - 1;
    while (r >= 0 && sv[r] < thr)
      r--;
    rank = r + 1;
  }
  // rows of U into W's slot
  if (act) {
#pragma unroll
    for (int i = 0; i < 16; i++) {
      W[j * n + i] = ((flip >> i) & 1ull) ? -u0[i] : u0[i];
      W[j * n + 16 + i] = ((flip >> (16 + i)) & 1ull) ? -u1[i] : u1[i];
      if (32 + i < n)
        W[j * n + 32 + i] = ((flip >> (32 + i)) & 1ull) ? -u2[i] : u2[i];
    }
  }
  __syncthreads();
  if (!bad && j < rank) {
    const int cj = perm[j];
    double t = 0.0;
    for (int c = 0; c < n; c++)
      t += W[c * n + cj] * b[c];
    tmp[j] = (1.0 / sv[j]) * t;
  }
  __syncthreads();
  if (act) {
    double t = 0.0;
    for (int i = 0; i < rank; i++)
      t += V[j * n + perm[i]] * tmp[i];
    x[j] = t;
  }
  __syncthreads();
}

// LDS doubles of smap_hyb (one system): W [n^2] (the Gram staging before the solve, U's rows after it)
// | V [n^2] (acc, cd, cidx before the solve) | b, sv, tmp, x [n each] | perm [n] ints.
__host__ __device__ inline size_t smap_hyb_lds_doubles(int n)
{
  return 2 * (size_t)n * n + 4 * (size_t)n + ((size_t)n + 1) / 2;
}

// Persistent grid over (target, theta) items of targets [t0, t0 + nt), one system per wavefront, for
// 33 <= n <= 48: about 28 KB of LDS per wavefront for n = 41, against 40.7 KB in smap_multi.
__global__ __launch_bounds__(64) void smap_hyb(SmapParams S, const double* __restrict__ D, int t0, int nt)
{
  extern __shared__ double smem[];
  const int n = S.Ea + 1;
  const int nacc = n * (n + 1) / 2 + n;
  const int lane = threadIdx.x;
  double* W = smem;
  double* V = W + (size_t)n * n;
  double* bvec = V + (size_t)n * n;
  double* sv = bvec + n;
  double* tmp = sv + n;
  double* xsol = tmp + n;
  int* perm = (int*)(xsol + n);
  double* acc = V;
  double* cd = acc + nacc;
  int* cidx = (int*)(cd + SM_CH);
  const int items = nt * S.T;
  for (int i0 = blockIdx.x; i0 < items; i0 += gridDim.x) {
    // the Gram staging lives in W's slot (SM_CH (n + 1) <= n^2), the accumulators in V's
    const unsigned valid = form_group(S, D, t0, nt, i0, n, 1, W, bvec, nullptr, SM_CH, acc, cd, cidx, lane);
    jacobi_svd_solve_hyb(n, W, V, bvec, sv, tmp, xsol, perm, !(valid & 1u), lane);
    write_group(S, t0, i0, n, 1, valid, xsol, lane);
    __syncthreads();
  }
}

// Launches the S-map kernel for n = E_a + 1 unknowns over targets [t0, t0 + nt) at every theta:
// register rows for n <= 24, U rows in registers for 33 <= n <= 48, LDS rows otherwise.  A persistent grid of one-wavefront blocks, as many as
// stay resident.  Returns 0, or -22 when n is beyond 64 (a system needs a lane per unknown).
// ---------------------------------------------------------------------------------------------
// Means of the valid distances of the targets of a dense block, each summed in library order (the sum
// form_group makes for an all-neighbour item).  A block takes DM_ROWS targets; their rows go through LDS
// DM_COLS columns at a time (coalesced; the next tile's loads are in flight while lane r adds row r's
// entries of the current one, one by one).
constexpr int DM_ROWS = 8, DM_COLS = 512, DM_THREADS = 256, DM_PER = DM_ROWS * DM_COLS / DM_THREADS;

__global__ __launch_bounds__(DM_THREADS) void dense_means(SmapParams S, const double* __restrict__ D, int t0,
                                                          int nt, double* __restrict__ mean)
{
  __shared__ double tile[DM_ROWS * (DM_COLS + 1) + 32];
  const int tid = threadIdx.x, r0 = blockIdx.x * DM_ROWS;
  const int rows = min(DM_ROWS, nt - r0);
  auto load = [&](int c0, double (&v)[DM_PER]) {
#pragma unroll
    for (int u = 0; u < DM_PER; u++) {
      const int e = tid + u * DM_THREADS, r = e / DM_COLS, c = c0 + e - r * DM_COLS;
      v[u] = r < rows && c < S.NL ? D[(size_t)(r0 + r) * S.NL + c] : -1.0;
    }
  };
  double v[DM_PER];
  load(0, v);
  double sum = 0.0;
  for (int c0 = 0; c0 < S.NL; c0 += DM_COLS) {
    // an entry that is not a neighbour (negative) enters the tile as +0.0: adding it leaves the sum as
    // it is (the sum starts at +0.0 and so is never -0.0), and the chain of adds needs no selects
#pragma unroll
    for (int u = 0; u < DM_PER; u++) {
      const int e = tid + u * DM_THREADS, r = e / DM_COLS, c = e - r * DM_COLS;
      tile[r * (DM_COLS + 1) + c] = v[u] >= 0.0 ? v[u] : 0.0;
    }
    __syncthreads();
    if (c0 + DM_COLS < S.NL)
      load(c0 + DM_COLS, v);
    if (tid < rows) {
      // the reads of the next 16 entries are in flight while the current 16 are added (the tile's tail
      // pads the last reads)
      const double* tr = tile + tid * (DM_COLS + 1);
      double a[16], b[16];
#pragma unroll
      for (int u = 0; u < 16; u++)
        a[u] = tr[u];
      for (int c = 0; c < DM_COLS; c += 32) {
#pragma unroll
        for (int u = 0; u < 16; u++)
          b[u] = tr[c + 16 + u];
#pragma unroll
        for (int u = 0; u < 16; u++)
          sum += a[u];
#pragma unroll
        for (int u = 0; u < 16; u++)
          a[u] = tr[c + 32 + u];
#pragma unroll
        for (int u = 0; u < 16; u++)
          sum += b[u];
      }
    }
    __syncthreads();
  }
  if (tid < rows)
    mean[r0 + tid] = sum / (double)S.nn_cnt[t0 + r0 + tid];
}

// X^T X and X^T y of the all-neighbour items of a dense block (nn_k = -1: every valid library point, in
// library order), for n <= 64.  One wavefront takes GI consecutive (target, theta) items (gd_items), so
// that every lane owns PER fixed entries of them over the whole library and keeps their sums in
// registers.  Library points go by in chunks of 64, one per lane: each lane weights its point for every
// item and stores the weighted row [w, w x_1 .. w x_E, w y] coordinate-major in LDS (zeros for a point
// that is not a neighbour); then every lane adds the chunk's 64 products to each of its entries, in
// library order.  Entry (p, q) is the sum over the neighbours of (w x_p)(w x_q), w = exp(-theta d /
// mean(d)), with the products and sums of form_group in its order, so the solve sees the same bits: a
// zero row adds +0.0, which leaves a sum unchanged (it starts at +0.0 and so is never -0.0).
// Output gw[item] = [W (n x n, both triangles) | b (n)].
constexpr int GD_MAXG = 16; // items a wavefront takes at most
constexpr int GD_RS = 66;   // LDS row stride (doubles): paired reads of up to 16 rows fall in different banks

__host__ __device__ constexpr int gd_nacc(int n)
{
  return n * (n + 1) / 2 + n;
}
__host__ __device__ constexpr int gd_items(int n)
{
  return gd_nacc(n) >= 64 ? 1 : (64 / gd_nacc(n) < GD_MAXG ? 64 / gd_nacc(n) : GD_MAXG);
}
__host__ __device__ constexpr int gd_per(int n)
{
  return (gd_items(n) * gd_nacc(n) + 63) / 64;
}

// Stores the lane's weighted row [w, w x_1 .. w x_E, w y] for each of the GI items; with SEL, zeros for
// an item whose distance d is negative (not a neighbour: its row may hold NaNs).  mrow is the point's
// row of the library (row 0 for a lane past its end).
template <int GI, bool SEL>
__device__ __forceinline__ void gd_stage(double* smem, int n, int lane, const double (&w)[GI], const double (&d)[GI],
                                         const double (&m0)[8], const double* mrow, double y)
{
  const int Ea = n - 1;
#pragma unroll
  for (int g = 0; g < GI; g++)
    smem[(g * (n + 1)) * GD_RS + lane] = !SEL || d[g] >= 0.0 ? w[g] : 0.0;
  for (int j0 = 0; j0 < Ea; j0 += 8) {
    // coordinates 0 .. 7 come fetched (m0); the others are read here
    double m[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
      m[u] = j0 == 0 ? m0[u] : mrow[min(j0 + u, Ea - 1)]; // in bounds: a row of the library
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (j0 + u < Ea)
#pragma unroll
        for (int g = 0; g < GI; g++)
          smem[(g * (n + 1) + 1 + j0 + u) * GD_RS + lane] = !SEL || d[g] >= 0.0 ? w[g] * m[u] : 0.0;
  }
#pragma unroll
  for (int g = 0; g < GI; g++)
    smem[(g * (n + 1) + n) * GD_RS + lane] = !SEL || d[g] >= 0.0 ? w[g] * y : 0.0;
}

template <int PER, int GI>
__global__ __launch_bounds__(64) void gram_dense(SmapParams S, const double* __restrict__ D, int t0, int nt,
                                                 const double* __restrict__ mean, double* __restrict__ gw)
{
  extern __shared__ double smem[]; // [GI][n + 1][GD_RS]: rows 0 .. n - 1 weighted coordinates, row n weighted y
  const int n = S.Ea + 1, Ea = S.Ea, T = S.T, lane = threadIdx.x;
  const int ntri = n * (n + 1) / 2, nacc = ntri + n;
  __shared__ double s_mean[GD_MAXG], s_theta[GD_MAXG];
  __shared__ int s_row[GD_MAXG];
  // this lane's entries: item g, rows p and q of the staged chunk (q = n: X^T y); oa/ob their LDS rows
  int ent[PER], oa[PER], ob[PER];
#pragma unroll
  for (int u = 0; u < PER; u++) {
    ent[u] = -1;
    oa[u] = ob[u] = 0;
    const int e = lane + 64 * u;
    if (e < GI * nacc) {
      const int g = e / nacc, r = e - g * nacc;
      int p = 0, q = n;
      if (r < ntri) {
        int rem = r;
        while (rem >= n - p) {
          rem -= n - p;
          p++;
        }
        q = p + rem;
      } else {
        p = r - ntri;
      }
      ent[u] = g << 16 | p << 8 | q;
      oa[u] = (g * (n + 1) + p) * GD_RS;
      ob[u] = (g * (n + 1) + q) * GD_RS;
    }
  }
  const int items = nt * T;
  for (int i0 = blockIdx.x * GI; i0 < items; i0 += gridDim.x * GI) {
    // per item: its distance row (-1: no all-neighbour system), mean and theta
    if (lane < GI) {
      const int item = i0 + lane;
      int row = -1;
      if (item < items) {
        const int tl = item / T, th = item - (item / T) * T, t = t0 + tl;
        if (S.nn_k[t] < 0) {
          row = tl;
          s_mean[lane] = mean[tl];
          s_theta[lane] = S.thetas[th];
        }
      }
      s_row[lane] = row;
    }
    __syncthreads();
    double acc[PER];
#pragma unroll
    for (int u = 0; u < PER; u++)
      acc[u] = 0.0;
    // the inputs of this lane's point of the next chunk, fetched a chunk ahead: its distance to each
    // item's target (items of one target share the load), its first 8 coordinates and its target value
    double nd[GI], nm[8], ny;
    auto fetch = [&](int c0) {
      const int pos = c0 + lane;
      const bool inlib = pos < S.NL;
#pragma unroll
      for (int g = 0; g < GI; g++)
        nd[g] = (g == 0 || s_row[g] != s_row[g - 1]) && inlib && s_row[g] >= 0
                  ? D[(size_t)s_row[g] * S.NL + pos]
                  : -1.0;
      const double* mrow = S.M + (size_t)(inlib ? pos : 0) * Ea;
#pragma unroll
      for (int u = 0; u < 8; u++)
        nm[u] = mrow[min(u, Ea - 1)]; // a row of the library (row 0 past its end): in bounds
      ny = inlib ? S.targets[pos] : 0.0;
    };
    fetch(0);
    for (int c0 = 0; c0 < S.NL; c0 += 64) {
      const int pos = c0 + lane;
      const bool inlib = pos < S.NL;
      double d[GI], m[8], q[GI], w[GI];
#pragma unroll
      for (int g = 0; g < GI; g++)
        d[g] = nd[g];
#pragma unroll
      for (int u = 0; u < 8; u++)
        m[u] = nm[u];
      const double y = ny;
      if (c0 + 64 < S.NL)
        fetch(c0 + 64);
      // the weight of the point for each item (items of one target share the quotient d / mean)
#pragma unroll
      for (int g = 0; g < GI; g++) {
        if (g > 0 && s_row[g] == s_row[g - 1]) {
          d[g] = d[g - 1];
          q[g] = q[g - 1];
        } else {
          q[g] = fmax(d[g], 0.0) / s_mean[g];
        }
      }
      bool ok = inlib;
#pragma unroll
      for (int g = 0; g < GI; g++) {
        w[g] = fdlibm_exp(-s_theta[g] * q[g]);
        ok = ok && d[g] >= 0.0;
      }
      const double* mrow = S.M + (size_t)(inlib ? pos : 0) * Ea;
      if (__all(ok)) {
        gd_stage<GI, false>(smem, n, lane, w, d, m, mrow, y);
      } else {
        gd_stage<GI, true>(smem, n, lane, w, d, m, mrow, y);
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < PER; u++) {
        const double* xa = smem + oa[u];
        const double* xb = smem + ob[u];
        double a = acc[u];
#pragma unroll 16
        for (int s2 = 0; s2 < 64; s2++)
          a += xa[s2] * xb[s2];
        acc[u] = a;
      }
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int en = ent[u];
      if (en < 0)
        continue;
      const int g = en >> 16, p = (en >> 8) & 255, q = en & 255;
      if (s_row[g] < 0)
        continue;
      double* out = gw + (size_t)(i0 + g) * (n * n + n);
      if (q < n) {
        out[p * n + q] = acc[u];
        out[q * n + p] = acc[u];
      } else {
        out[n * n + p] = acc[u];
      }
    }
    __syncthreads();
  }
}

// Forms the all-neighbour systems of targets t0 .. t0 + nt - 1 into gw (the scratch gram_dense_bytes
// gives, whose tail holds the means); the solve kernels then copy them (S.gram = gw).
size_t gram_dense_bytes(int n, int nt, int T)
{
  return sizeof(double) * ((size_t)nt * T * ((size_t)n * n + n) + (size_t)nt);
}

int gram_dense_launch(int n, hipStream_t s, const SmapParams& S, const double* D, int t0, int nt, double* gw)
{
  if (n > 64 || n < 2)
    return -22;
  double* mean = gw + (size_t)nt * S.T * ((size_t)n * n + n);
  hipLaunchKernelGGL(dense_means, dim3((nt + DM_ROWS - 1) / DM_ROWS), dim3(DM_THREADS), 0, s, S, D, t0, nt, mean);
  hipError_t e = hipGetLastError();
  const int GI = gd_items(n), per = gd_per(n);
  const void* fn = GI == 12   ? (const void*)gram_dense<1, 12>
                   : GI == 7  ? (const void*)gram_dense<1, 7>
                   : GI == 4  ? (const void*)gram_dense<1, 4>
                   : GI == 3  ? (const void*)gram_dense<1, 3>
                   : GI == 2  ? (const void*)gram_dense<1, 2>
                   : GI != 1  ? nullptr
                   : per <= 1 ? (const void*)gram_dense<1, 1>
                   : per <= 2 ? (const void*)gram_dense<2, 1>
                   : per <= 3 ? (const void*)gram_dense<3, 1>
                   : per <= 6 ? (const void*)gram_dense<6, 1>
                   : per <= 12 ? (const void*)gram_dense<12, 1>
                               : (const void*)gram_dense<36, 1>;
  if (!fn)
    return -22;
  const size_t lds = sizeof(double) * (size_t)GI * (n + 1) * GD_RS;
  const long long groups = ((long long)nt * S.T + GI - 1) / GI;
  int dev = 0, ncu = 0, occ = 0;
  if (e == hipSuccess)
    e = hipGetDevice(&dev);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 64, lds);
  if (e != hipSuccess)
    return -(int)e - 1000;
  const int grid = (int)std::max<long long>(1, std::min<long long>(groups, (long long)ncu * std::max(occ, 1)));
  const double* cmean = mean;
  void* args[] = { (void*)&S, (void*)&D, (void*)&t0, (void*)&nt, (void*)&cmean, (void*)&gw };
  e = hipLaunchKernel(fn, dim3(grid), dim3(64), args, lds, s);
  return e == hipSuccess ? 0 : -(int)e - 1000;
}

// ---------------------------------------------------------------------------------------------
// Wide systems (65 <= n <= SMAP_BIG_MAX unknowns): one (target, theta) system per 256-thread block, a
// grid-stride loop over the items.  W lives in LDS; U and V live transposed in the block's slice of
// global scratch (column i of U at Ut[i n ..], so a rotation's two column updates are contiguous
// rows).  The Gram entries are summed neighbour by neighbour as in form_group, and the solve makes
// jacobi_svd_solve_wave's operations in its order, with threads in place of lanes and a barrier
// where the wavefront version orders its LDS traffic.
constexpr int SB_THREADS = 256;
constexpr int SB_CH = 8; // neighbour rows staged per Gram chunk

__host__ __device__ inline size_t smap_big_lds_doubles(int n)
{
  // W [n^2] | b, sv, tmp, x [n each] | Xc [SB_CH n] | yc, cd, wts [SB_CH each] | cidx [SB_CH ints] | scalars
  return (size_t)n * n + 4 * (size_t)n + (size_t)SB_CH * n + 4 * SB_CH + 8;
}

__device__ __forceinline__ double block_max(double v, double* red)
{
  v = wave_max(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
    red[w] = v;
  __syncthreads();
  double m = red[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); i++)
    m = fmax(m, red[i]);
  return m;
}

__global__ __launch_bounds__(SB_THREADS) void smap_big(SmapParams S, const double* __restrict__ D, int t0, int nt)
{
  extern __shared__ double smem[];
  __shared__ double red[SB_THREADS / 64];
  const int n = S.Ea + 1, Ea = S.Ea, T = S.T, tid = threadIdx.x;
  double* W = smem;
  double* bv = W + (size_t)n * n;
  double* sv = bv + n;
  double* tmp = sv + n;
  double* xs = tmp + n;
  double* Xc = xs + n;
  double* yc = Xc + (size_t)SB_CH * n;
  double* cd = yc + SB_CH;
  double* wts = cd + SB_CH;
  int* cidx = (int*)(wts + SB_CH);
  double* scal = wts + SB_CH + SB_CH / 2; // [0] mean, [1] chunk count, [2] sort position, [3] nonzero
  double* Ut = S.big + (size_t)blockIdx.x * 2 * n * n;
  double* Vt = Ut + (size_t)n * n;
  const int items = nt * T;
  for (int item = blockIdx.x; item < items; item += gridDim.x) {
    const int tl = item / T, th = item - (item / T) * T;
    const int t = t0 + tl;
    if (t >= S.NP)
      continue;
    const int kk = S.nn_k[t];
    if (kk == 0)
      continue;
    const bool dense = kk < 0;
    const int k = dense ? S.nn_cnt[t] : kk;
    const double* drow = dense ? D + (size_t)tl * S.NL : nullptr;
    const int* li = S.nn_i + (size_t)(t - S.nn_t0) * S.nn_stride;
    const double* lsd = S.nn_s + (size_t)(t - S.nn_t0) * S.nn_stride;
    __syncthreads();
    if (tid == 0) {
      // mean of the selected distances, summed in neighbour order (Eigen mean = sum / size)
      double sum = 0.0;
      if (dense) {
        for (int i = 0; i < S.NL; i++)
          if (drow[i] >= 0.0)
            sum += drow[i];
      } else {
        for (int j = 0; j < k; j++)
          sum += lsd[j];
      }
      scal[0] = sum / (double)k;
    }
    for (int e = tid; e < n * n; e += SB_THREADS)
      W[e] = 0.0;
    for (int e = tid; e < n; e += SB_THREADS)
      bv[e] = 0.0;
    __syncthreads();
    const double mean = scal[0];
    const double theta = S.thetas[th];
    // X^T X (entries p <= q, accumulated in W[p n + q]) and X^T y, neighbour by neighbour
    int pos = 0, done = 0;
    while (done < k) {
      if (tid == 0) {
        int c = 0;
        if (dense) {
          while (c < SB_CH && pos < S.NL) {
            if (drow[pos] >= 0.0) {
              cidx[c] = pos;
              cd[c] = drow[pos];
              c++;
            }
            pos++;
          }
        } else {
          c = min(SB_CH, k - done);
          for (int j = 0; j < c; j++) {
            cidx[j] = li[done + j];
            cd[j] = lsd[done + j];
          }
        }
        scal[1] = (double)c;
      }
      __syncthreads();
      const int cnum = (int)scal[1];
      if (cnum == 0)
        break;
      if (tid < cnum) {
        const double w = fdlibm_exp(-theta * (cd[tid] / mean));
        wts[tid] = w;
        yc[tid] = w * S.targets[cidx[tid]];
      }
      __syncthreads();
      for (int e = tid; e < cnum * n; e += SB_THREADS) {
        const int c = e / n, j = e - (e / n) * n;
        const double w = wts[c];
        Xc[e] = j == 0 ? w : w * S.M[(size_t)cidx[c] * Ea + (j - 1)];
      }
      __syncthreads();
      for (int e = tid; e < n * n + n; e += SB_THREADS) {
        if (e < n * n) {
          const int p = e / n, q = e - (e / n) * n;
          if (q < p)
            continue;
          double a = W[e];
          for (int c = 0; c < cnum; c++)
            a += Xc[c * n + p] * Xc[c * n + q];
          W[e] = a;
        } else {
          const int p = e - n * n;
          double a = bv[p];
          for (int c = 0; c < cnum; c++)
            a += Xc[c * n + p] * yc[c];
          bv[p] = a;
        }
      }
      __syncthreads();
      done += cnum;
    }
    for (int e = tid; e < n * n; e += SB_THREADS) {
      const int p = e / n, q = e - (e / n) * n;
      if (q < p)
        W[e] = W[q * n + p];
    }
    __syncthreads();

    // Jacobi SVD solve (jacobi_svd_solve_wave's steps)
    double cm = 0.0;
    for (int e = tid; e < n * n; e += SB_THREADS)
      cm = fmax(cm, fabs(W[e]));
    double scale = block_max(cm, red);
    const bool bad = !isfinite(scale);
    if (scale == 0.0)
      scale = 1.0;
    if (!bad) {
      for (int e = tid; e < n * n; e += SB_THREADS) {
        W[e] = W[e] / scale;
        const int r = e / n, c = e - (e / n) * n;
        Ut[e] = (r == c) ? 1.0 : 0.0;
        Vt[e] = (r == c) ? 1.0 : 0.0;
      }
    }
    __syncthreads();
    int rank = 0;
    if (!bad) {
      const double precision = 2.0 * DBL_EPSILON;
      double maxDiag = 0.0;
      for (int i = 0; i < n; i++)
        maxDiag = fmax(maxDiag, fabs(W[i * n + i]));
      bool finished = false;
      int sweeps = 0;
      while (!finished && sweeps < 1000) {
        finished = true;
        sweeps++;
        for (int p = 1; p < n; p++) {
          for (int q = 0; q < p; q++) {
            const double thr = fmax(DBL_MIN, precision * maxDiag);
            const double wpq = W[p * n + q], wqp = W[q * n + p];
            if (fabs(wpq) > thr || fabs(wqp) > thr) {
              finished = false;
              Rot jl, jr;
              real_2x2(W[p * n + p], wpq, wqp, W[q * n + q], &jl, &jr);
              __syncthreads();
              if (!(jl.c == 1.0 && jl.s == 0.0))
                for (int j = tid; j < n; j += SB_THREADS) {
                  const double xv = W[p * n + j], yv = W[q * n + j];
                  W[p * n + j] = jl.c * xv + jl.s * yv;
                  W[q * n + j] = -jl.s * xv + jl.c * yv;
                  const double xu = Ut[p * n + j], yu = Ut[q * n + j];
                  Ut[p * n + j] = jl.c * xu + jl.s * yu;
                  Ut[q * n + j] = -jl.s * xu + jl.c * yu;
                }
              __syncthreads();
              const double tc = jr.c, ts = -jr.s;
              if (!(tc == 1.0 && ts == 0.0))
                for (int j = tid; j < n; j += SB_THREADS) {
                  const double xv = W[j * n + p], yv = W[j * n + q];
                  W[j * n + p] = tc * xv + ts * yv;
                  W[j * n + q] = -ts * xv + tc * yv;
                  const double xw = Vt[p * n + j], yw = Vt[q * n + j];
                  Vt[p * n + j] = tc * xw + ts * yw;
                  Vt[q * n + j] = -ts * xw + tc * yw;
                }
              __syncthreads();
              maxDiag = fmax(maxDiag, fmax(fabs(W[p * n + p]), fabs(W[q * n + q])));
            }
          }
        }
      }
      // non-negative singular values
      for (int i = 0; i < n; i++) {
        const double a = W[i * n + i];
        if (a < 0.0)
          for (int j = tid; j < n; j += SB_THREADS)
            Ut[i * n + j] = -Ut[i * n + j];
        if (tid == 0)
          sv[i] = fabs(a) * scale;
      }
      __syncthreads();
      // decreasing order, first maximum wins
      int nonzero = n;
      for (int i = 0; i < n; i++) {
        int sp = i;
        double mx = sv[i];
        for (int j = i + 1; j < n; j++)
          if (sv[j] > mx) {
            mx = sv[j];
            sp = j;
          }
        if (mx == 0.0) {
          nonzero = i;
          break;
        }
        if (sp != i) {
          __syncthreads();
          if (tid == 0) {
            const double a = sv[i];
            sv[i] = sv[sp];
            sv[sp] = a;
          }
          for (int j = tid; j < n; j += SB_THREADS) {
            double a = Ut[i * n + j];
            Ut[i * n + j] = Ut[sp * n + j];
            Ut[sp * n + j] = a;
            a = Vt[i * n + j];
            Vt[i * n + j] = Vt[sp * n + j];
            Vt[sp * n + j] = a;
          }
          __syncthreads();
        }
      }
      const double thr = fmax(sv[0] * ((double)n * DBL_EPSILON), DBL_MIN);
      int r = nonzero - 1;
      while (r >= 0 && sv[r] < thr)
        r--;
      rank = r + 1;
      for (int i = tid; i < rank; i += SB_THREADS) {
        double s = 0.0;
        for (int j = 0; j < n; j++)
          s += Ut[i * n + j] * bv[j];
        tmp[i] = (1.0 / sv[i]) * s;
      }
      __syncthreads();
    }
    for (int l = tid; l < n; l += SB_THREADS) {
      double s = 0.0;
      for (int i = 0; i < rank; i++)
        s += Vt[i * n + l] * tmp[i];
      xs[l] = s;
    }
    __syncthreads();
    if (tid == 0) {
      double r = xs[0];
      const double* yp = S.Mp + (size_t)t * Ea;
      for (int c = 0; c < Ea; c++)
        if (yp[c] != MISSING_D)
          r += yp[c] * xs[c + 1];
      S.pred[(size_t)th * S.NP + t] = r;
      S.rc[(size_t)th * S.NP + t] = RC_SUCCESS;
    }
    if (S.coeffs && th == T - 1)
      for (int j = tid; j < n; j += SB_THREADS) {
        const double c = xs[j];
        S.coeffs[(size_t)t * n + j] = fabs(c) < 1.0e-11 ? MISSING_D : c;
      }
  }
}

// smap_pack over groups of G systems: the largest G whose LDS leaves the blocks per CU that the kernel's
// registers allow (or, with max_group, the largest G that fits).
static int smap_pack_launch(int n, hipStream_t s, const SmapParams& S, const double* D, int t0, int nt, bool max_group)
{
  const void* fn = n <= 8    ? (const void*)smap_pack<8, 0>
                   : n <= 16 ? (const void*)smap_pack<16, 0>
                   : n <= 17 ? (const void*)smap_pack<16, 1>
                   : n <= 20 ? (const void*)smap_pack<16, 4>
                             : (const void*)smap_pack<16, 8>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  int dev = 0, ncu = 0, occ1 = 0, occ = 0;
  if (e == hipSuccess)
    e = hipGetDevice(&dev);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ1, fn, 64, sizeof(double) * smap_pack_lds_doubles(n, 1));
  int G = smap_pack_group(n);
  while (e == hipSuccess) {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 64, sizeof(double) * smap_pack_lds_doubles(n, G));
    if (e != hipSuccess || G == 1 || (max_group ? occ >= 1 : occ >= occ1))
      break;
    G--;
  }
  if (e != hipSuccess)
    return -(int)e - 1000;
  const size_t lds = sizeof(double) * smap_pack_lds_doubles(n, G);
  const long long groups = ((long long)nt * S.T + G - 1) / G;
  const int grid = (int)std::max<long long>(1, std::min<long long>(groups, (long long)ncu * std::max(occ, 1)));
  void* args[] = { (void*)&S, (void*)&D, (void*)&t0, (void*)&nt, (void*)&G };
  e = hipLaunchKernel(fn, dim3(grid), dim3(64), args, lds, s);
  return e == hipSuccess ? 0 : -(int)e - 1000;
}

int smap_launch(int n, hipStream_t s, const SmapParams& S, const double* D, int t0, int nt, bool lds_rows,
                int solve_mode)
{
  if (solve_mode == SMAP_SOLVE_AUTO)
    solve_mode = n >= 9 && n <= 24 && (long long)nt * S.T >= SMAP_PACK_MIN_ITEMS ? SMAP_SOLVE_PACK : SMAP_SOLVE_ROWS;
  if (n > 64 && n <= SMAP_BIG_MAX) {
    if (!S.big || S.big_blocks < 1)
      return -22;
    const size_t lds = sizeof(double) * smap_big_lds_doubles(n);
    hipError_t e = hipFuncSetAttribute((const void*)smap_big, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const long long items = (long long)nt * S.T;
    const int grid = (int)std::max<long long>(1, std::min<long long>(items, S.big_blocks));
    void* args[] = { (void*)&S, (void*)&D, (void*)&t0, (void*)&nt };
    if (e == hipSuccess)
      e = hipLaunchKernel((const void*)smap_big, dim3(grid), dim3(SB_THREADS), args, lds, s);
    return e == hipSuccess ? 0 : -(int)e - 1000;
  }
  if (n > 64 || n < 2)
    return -22;
  if (!lds_rows && n >= 2 && n <= 24 && (solve_mode == SMAP_SOLVE_PACK || solve_mode == SMAP_SOLVE_PACK_MAX))
    return smap_pack_launch(n, s, S, D, t0, nt, solve_mode == SMAP_SOLVE_PACK_MAX);
  const int G = smap_group(n);
  const void* fn = (const void*)smap_multi;
  size_t lds = sizeof(double) * smap_lds_doubles(n, G);
  int waves = 1; // wavefronts per block (each with G systems)
  if (!lds_rows && n <= 24 && solve_mode == SMAP_SOLVE_BLOCK) {
    constexpr int WB = BLK_WAVES;
    fn = n <= 8    ? (const void*)smap_blk<8, 0, WB>
         : n <= 16 ? (const void*)smap_blk<16, 0, WB>
         : n <= 17 ? (const void*)smap_blk<16, 1, WB>
         : n <= 20 ? (const void*)smap_blk<16, 4, WB>
                   : (const void*)smap_blk<16, 8, WB>;
    lds = sizeof(double) * smap_blk_lds_doubles(n, G, WB);
    waves = WB;
  } else if (!lds_rows && n <= 24) {
    fn = n <= 8    ? (const void*)smap_rows<8, 0>
         : n <= 16 ? (const void*)smap_rows<16, 0>
         : n <= 17 ? (const void*)smap_rows<16, 1>
         : n <= 20 ? (const void*)smap_rows<16, 4>
                   : (const void*)smap_rows<16, 8>;
    lds = sizeof(double) * smap_rows_lds_doubles(n, G);
  } else if (!lds_rows && n >= 33 && n <= 48) {
    fn = (const void*)smap_hyb;
    lds = sizeof(double) * smap_hyb_lds_doubles(n);
  }
  // HIP errors come back as -(error) - 1000, as elsewhere in the library
  hipError_t e = hipSuccess;
  if (lds > 64 * 1024)
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  const long long groups = ((long long)nt * S.T + G * waves - 1) / (G * waves);
  int dev = 0, ncu = 0, occ = 0;
  if (e == hipSuccess)
    e = hipGetDevice(&dev);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 64 * waves, lds);
  if (e != hipSuccess)
    return -(int)e - 1000;
  const int grid = (int)std::max<long long>(1, std::min<long long>(groups, (long long)ncu * std::max(occ, 1)));
  void* args[] = { (void*)&S, (void*)&D, (void*)&t0, (void*)&nt };
  e = hipLaunchKernel(fn, dim3(grid), dim3(64 * waves), args, lds, s);
  return e == hipSuccess ? 0 : -(int)e - 1000;
}

} // namespace fedm
