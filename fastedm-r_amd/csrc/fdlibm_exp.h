// exp(x) by FDLIBM's published method (Sun Microsystems e_exp.c, 1993): argument reduction
// x = k ln2 + r with a two-part ln2, a degree-5 rational-polynomial approximation of exp(r), and an
// exact scaling by 2^k.  Used for the simplex and S-map weights on the GPU so that the FP64 weights
// are reproducible bit for bit by the CPU oracle, which evaluates the same expression.  Accuracy is
// within 1 ulp, as for the libm exp the reference calls.
#pragma once
#include <hip/hip_runtime.h>

namespace fedm {

__device__ __forceinline__ double fdlibm_exp(double x)
{
  const double ln2HI = 6.93147180369123816490e-01, ln2LO = 1.90821492927058770002e-10,
               invln2 = 1.44269504088896338700e+00;
  const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03, P3 = 6.61375632143793436117e-05,
               P4 = -1.65339022054652515390e-06, P5 = 4.13813679705723846039e-08;
  if (x != x)
    return x;
  if (x > 709.782712893383973096)
    return __builtin_huge_val();
  if (x < -745.13321910194110842)
    return 0.0;
  const double ax = fabs(x);
  if (ax < 3.725290298461914e-09) // |x| < 2^-28
    return 1.0 + x;
  int k;
  double hi, lo;
  if (ax > 0.5 * 0.6931471805599453) {
    if (ax < 1.5 * 0.6931471805599453) {
      k = x < 0 ? -1 : 1;
      hi = x - k * ln2HI;
      lo = k * ln2LO;
    } else {
      k = (int)(invln2 * x + (x < 0 ? -0.5 : 0.5));
      const double t = k;
      hi = x - t * ln2HI;
      lo = t * ln2LO;
    }
  } else {
    k = 0;
    hi = x;
    lo = 0.0;
  }
  const double r = hi - lo;
  const double t = r * r;
  const double c = r - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
  if (k == 0)
    return 1.0 - ((r * c) / (c - 2.0) - r);
  const double y = 1.0 - ((lo - (r * c) / (2.0 - c)) - hi);
  return ldexp(y, k);
}

} // namespace fedm
