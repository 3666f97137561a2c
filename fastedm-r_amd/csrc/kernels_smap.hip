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

#include "kernels_common.h"
#include "fdlibm_exp.h"

namespace fedm {

constexpr int SM_CH = 32; // neighbours staged per Gram chunk

__host__ __device__ inline int smap_group(int n)
{
  return n > 64 ? 0 : (64 / n > 8 ? 8 : 64 / n);
}

// LDS doubles: W [G n^2] | U, V [G n^2 each] (the Gram staging Xc [SM_CH n], yc [SM_CH] lives there
// before the solve starts) | b, sv, tmp, x [G n each] | acc [n (n+1)/2 + n] | cd [SM_CH] | cidx [SM_CH]
// ints.
__host__ __device__ inline size_t smap_lds_doubles(int n, int G)
{
  const size_t nacc = (size_t)n * (n + 1) / 2 + n;
  const size_t uv = std::max((size_t)2 * G * n * n, (size_t)SM_CH * n + SM_CH);
  return (size_t)G * n * n + uv + 4 * (size_t)G * n + nacc + SM_CH + SM_CH / 2;
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
  // non-negative singular values
  for (int i = 0; i < n; i++) {
    const double a = bad ? 0.0 : Wg[i * n + i];
    if (!bad && a < 0.0)
      Ug[j * n + i] = -Ug[j * n + i];
    if (!bad && j == 0)
      svg[i] = fabs(a) * scale;
  }
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

// Persistent grid over groups of G (target, theta) items of targets [t0, t0 + nt).
__global__ __launch_bounds__(64) void smap_multi(SmapParams S, const double* __restrict__ D, int t0, int nt)
{
  extern __shared__ double smem[];
  const int Ea = S.Ea, n = Ea + 1, T = S.T;
  const int G = smap_group(n);
  const int ntri = n * (n + 1) / 2, nacc = ntri + n;
  const int lane = threadIdx.x;
  double* W = smem;
  double* U = W + (size_t)G * n * n;
  double* V = U + (size_t)G * n * n;
  const size_t uv = max((size_t)2 * G * n * n, (size_t)SM_CH * n + SM_CH);
  double* bvec = U + uv;
  double* sv = bvec + G * n;
  double* tmp = sv + G * n;
  double* xsol = tmp + G * n;
  double* acc = xsol + G * n;
  double* cd = acc + nacc;
  int* cidx = (int*)(cd + SM_CH);
  double* Xc = U;
  double* yc = Xc + SM_CH * n;

  const int items = nt * T;
  for (int i0 = blockIdx.x * G; i0 < items; i0 += gridDim.x * G) {
    // Gram systems of the group's items, one after the other
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
      const bool dense = kk < 0;
      const int k = dense ? S.nn_cnt[t] : kk;
      const double* drow = dense ? D + (size_t)tl * S.NL : nullptr;
      const int* li = S.nn_i + (size_t)(t - S.nn_t0) * S.nn_stride;
      const double* lsd = S.nn_s + (size_t)(t - S.nn_t0) * S.nn_stride;
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
            const bool v = lane < SM_CH && i < S.NL && drow[i] >= 0.0;
            const unsigned long long bb = __ballot(v);
            const int before = __popcll(bb & ((1ull << lane) - 1ull));
            if (v) {
              cidx[before] = i;
              cd[before] = drow[i];
            }
            cnum = __popcll(bb);
            pos += SM_CH;
          }
        } else {
          cnum = min(SM_CH, k - done);
          if (lane < cnum) {
            cidx[lane] = li[done + lane];
            cd[lane] = lsd[done + lane];
          }
        }
        if (cnum == 0)
          break;
        __syncthreads();
        if (lane < cnum) {
          const int ci = cidx[lane];
          const double w = fdlibm_exp(-theta * (cd[lane] / mean));
          const double* row = S.M + (size_t)ci * Ea;
          Xc[lane * n] = w;
          for (int j = 1; j < n; j++)
            Xc[lane * n + j] = w * row[j - 1];
          yc[lane] = w * S.targets[ci];
        }
        __syncthreads();
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
        __syncthreads();
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
      __syncthreads();
    }

    const int gl = lane / n;
    const bool mine = gl < G && ((valid >> gl) & 1u);
    jacobi_svd_solve_multi(n, G, W, U, V, bvec, sv, tmp, xsol, !mine, lane);

    if (mine) {
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
    __syncthreads();
  }
}

// LDS bytes of smap_multi for n unknowns.
size_t smap_lds_bytes(int n)
{
  return sizeof(double) * smap_lds_doubles(n, smap_group(n));
}

} // namespace fedm
