// Two-sided Jacobi SVD solve of a small square system on one wavefront, with the matrices in LDS.
//
// This follows the published algorithm of Eigen 3.4's JacobiSVD for a square matrix with thin U/V
// (the reference solves its S-map normal equations with JacobiSVD<MatrixXd>(X^T X).solve(X^T y),
// src/edm.cpp:780-782): scale by the largest |a_ij|; cyclic sweeps over p = 1..n-1, q = 0..p-1 while
// an off-diagonal pair exceeds max(DBL_MIN, 2 eps * max|diag|); each 2x2 block made symmetric by a
// plane rotation then diagonalised by a Jacobi rotation; singular values made non-negative, rescaled
// and sorted in decreasing order; the solve keeps singular values >= max(s_0 * n * eps, DBL_MIN).
// Row rotations run one column per lane, column rotations one row per lane; the scalar 2x2 work is
// done redundantly by every lane from LDS values, so control flow stays uniform.
#pragma once

namespace fedm {

struct Rot
{
  double c, s;
};

__device__ __forceinline__ Rot make_jacobi(double x, double y, double z)
{
  const double deno = 2.0 * fabs(y);
  if (deno < DBL_MIN)
    return { 1.0, 0.0 };
  const double tau = (x - z) / deno;
  const double w = sqrt(tau * tau + 1.0);
  const double t = 1.0 / (tau > 0.0 ? tau + w : tau - w);
  const double sign_t = t > 0.0 ? 1.0 : -1.0;
  const double nrm = 1.0 / sqrt(t * t + 1.0);
  // y / |y| is exactly +-1 for finite nonzero y
  const double s = -sign_t * copysign(1.0, y) * fabs(t) * nrm;
  return { nrm, s };
}

__device__ __forceinline__ void real_2x2(double m00, double m01, double m10, double m11, Rot* jl, Rot* jr)
{
  Rot r1;
  const double t = m00 + m11;
  const double d = m10 - m01;
  if (fabs(d) < DBL_MIN) {
    r1 = { 1.0, 0.0 };
  } else {
    const double u = t / d;
    const double tmp = sqrt(1.0 + u * u);
    r1 = { u / tmp, 1.0 / tmp };
  }
  if (!(r1.c == 1.0 && r1.s == 0.0)) {
    const double a0 = m00, a1 = m01, b0 = m10, b1 = m11;
    m00 = r1.c * a0 + r1.s * b0;
    m01 = r1.c * a1 + r1.s * b1;
    m11 = -r1.s * a1 + r1.c * b1;
  }
  *jr = make_jacobi(m00, m01, m11);
  const Rot jt{ jr->c, -jr->s };
  jl->c = r1.c * jt.c - r1.s * jt.s;
  jl->s = r1.c * jt.s + r1.s * jt.c;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v = fmax(v, __shfl_xor(v, off));
  return v;
}

// W (n x n, row-major, stride n) is overwritten. Result in x[0..n). U, V, sv, tmp are scratch (LDS).
__device__ void jacobi_svd_solve_wave(int n, double* W, double* U, double* V, const double* b, double* sv,
                                      double* tmp, double* x, int lane)
{
  const int nn = n * n;
  double scale = 0.0;
  for (int e = lane; e < nn; e += 64)
    scale = fmax(scale, fabs(W[e]));
  scale = wave_max(scale);
  if (!isfinite(scale)) {
    if (lane < n)
      x[lane] = 0.0;
    __syncthreads();
    return;
  }
  if (scale == 0.0)
    scale = 1.0;
  for (int e = lane; e < nn; e += 64) {
    W[e] = W[e] / scale;
    const int r = e / n, c = e - r * n;
    U[e] = (r == c) ? 1.0 : 0.0;
    V[e] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
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
          // rows p, q of W by jl
          if (lane < n && !(jl.c == 1.0 && jl.s == 0.0)) {
            const double xv = W[p * n + lane], yv = W[q * n + lane];
            W[p * n + lane] = jl.c * xv + jl.s * yv;
            W[q * n + lane] = -jl.s * xv + jl.c * yv;
          }
          // columns p, q of U by jl (U.applyOnTheRight(p, q, jl^T))
          if (lane < n && !(jl.c == 1.0 && jl.s == 0.0)) {
            const double xv = U[lane * n + p], yv = U[lane * n + q];
            U[lane * n + p] = jl.c * xv + jl.s * yv;
            U[lane * n + q] = -jl.s * xv + jl.c * yv;
          }
          __syncthreads();
          // columns p, q of W and V by jr (applyOnTheRight: rotation (c, -s))
          const double tc = jr.c, ts = -jr.s;
          if (lane < n && !(tc == 1.0 && ts == 0.0)) {
            const double xv = W[lane * n + p], yv = W[lane * n + q];
            W[lane * n + p] = tc * xv + ts * yv;
            W[lane * n + q] = -ts * xv + tc * yv;
            const double xw = V[lane * n + p], yw = V[lane * n + q];
            V[lane * n + p] = tc * xw + ts * yw;
            V[lane * n + q] = -ts * xw + tc * yw;
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
    if (a < 0.0 && lane < n)
      U[lane * n + i] = -U[lane * n + i];
    if (lane == 0)
      sv[i] = fabs(a) * scale;
  }
  __syncthreads();
  // decreasing order, first maximum wins
  int nonzero = n;
  for (int i = 0; i < n; i++) {
    int pos = i;
    double mx = sv[i];
    for (int j = i + 1; j < n; j++)
      if (sv[j] > mx) {
        mx = sv[j];
        pos = j;
      }
    if (mx == 0.0) {
      nonzero = i;
      break;
    }
    if (pos != i) {
      __syncthreads();
      if (lane == 0) {
        const double a = sv[i];
        sv[i] = sv[pos];
        sv[pos] = a;
      }
      if (lane < n) {
        double a = U[lane * n + i];
        U[lane * n + i] = U[lane * n + pos];
        U[lane * n + pos] = a;
        a = V[lane * n + i];
        V[lane * n + i] = V[lane * n + pos];
        V[lane * n + pos] = a;
      }
      __syncthreads();
    }
  }
  const double thr = fmax(sv[0] * ((double)n * DBL_EPSILON), DBL_MIN);
  int r = nonzero - 1;
  while (r >= 0 && sv[r] < thr)
    r--;
  const int rank = r + 1;
  if (lane < rank) {
    double s = 0.0;
    for (int j = 0; j < n; j++)
      s += U[j * n + lane] * b[j];
    tmp[lane] = (1.0 / sv[lane]) * s;
  }
  __syncthreads();
  if (lane < n) {
    double s = 0.0;
    for (int i = 0; i < rank; i++)
      s += V[lane * n + i] * tmp[i];
    x[lane] = s;
  }
  __syncthreads();
}

} // namespace fedm
