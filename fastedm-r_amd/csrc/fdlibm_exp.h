// exp(x) by FDLIBM's published method (Sun Microsystems e_exp.c, 1993): argument reduction
// x = k ln2 + r with a two-part ln2, a degree-5 rational-polynomial approximation of exp(r), and an
// exact scaling by 2^k.  Used for the simplex and S-map weights on the GPU so that the FP64 weights
// are reproducible bit for bit by the CPU oracle, which evaluates the same expression.  Accuracy is
// within 1 ulp, as for the libm exp the reference calls.
#pragma once
#include <hip/hip_runtime.h>

namespace fedm {

// Written without branches, so that the exps of several weights interleave: the reduction and the
// polynomial run for every x (clamped so that k stays in range), and the special cases (NaN, overflow,
// underflow, |x| < 2^-28) select their results at the end.  k = 0 needs no case of its own: with
// t = 0, hi = x and lo = +0, and 1 - ((lo - q) - hi) with q = (r c) / (2 - c) equals the k = 0 form
// 1 - ((r c) / (c - 2) - r) bit for bit (negating a divisor or subtracting from +0 is exact for the
// nonzero x that reach it).
__device__ __forceinline__ double fdlibm_exp(double x)
{
  const double ln2HI = 6.93147180369123816490e-01, ln2LO = 1.90821492927058770002e-10,
               invln2 = 1.44269504088896338700e+00;
  const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03, P3 = 6.61375632143793436117e-05,
               P4 = -1.65339022054652515390e-06, P5 = 4.13813679705723846039e-08;
  const double xc = fmin(fmax(x, -746.0), 710.0);
  const double axc = fabs(xc);
  const int kr = (int)(invln2 * xc + (xc < 0 ? -0.5 : 0.5));
  const int k = axc > 0.5 * 0.6931471805599453 ? (axc < 1.5 * 0.6931471805599453 ? (xc < 0 ? -1 : 1) : kr) : 0;
  const double t = k;
  const double hi = xc - t * ln2HI;
  const double lo = t * ln2LO;
  const double r = hi - lo;
  const double rr = r * r;
  const double c = r - rr * (P1 + rr * (P2 + rr * (P3 + rr * (P4 + rr * P5))));
  const double y = 1.0 - ((lo - (r * c) / (2.0 - c)) - hi);
  double res = ldexp(y, k);
  res = fabs(x) < 3.725290298461914e-09 ? 1.0 + x : res; // |x| < 2^-28
  res = x < -745.13321910194110842 ? 0.0 : res;
  res = x > 709.782712893383973096 ? __builtin_huge_val() : res;
  return x != x ? x : res;
}

} // namespace fedm
