// FP64 exp and log for the chemistry kernels, with a fixed, documented
// operation count (used by the FLOP accounting of the roofline).
//
// exp: Cody-Waite argument reduction (W. J. Cody, W. Waite, "Software Manual
//   for the Elementary Functions", 1980): x = k ln2 + r, k = rint(x/ln2),
//   ln2 split into a 32-bit head and a tail so that k*ln2_hi is exact for
//   |k| <= 1100; e^r on |r| <= ln2/2 by its degree-13 Taylor polynomial in
//   Horner form (truncation < 5e-18 relative); result = 2^k * p by ldexp.
//   Inputs below -746 give 0 (through ldexp underflow), above 710 give +inf,
//   NaN propagates.  Ops: 1 mul, 1 rint, 15 fma, 1 ldexp, 1 cvt (plus the
//   clamp: min, max and a NaN select).
// log: x = 2^e m with m in [sqrt(1/2), sqrt(2)) (frexp); with f = m - 1 and
//   s = f/(2 + f), ln m = 2 atanh(s) = 2s(1 + z/3 + z^2/5 + ...), z = s^2,
//   |s| <= 0.1716, truncated after z^10 (< 1e-18 relative); ln x =
//   e ln2_hi + (2s P(z) + e ln2_lo).  x = 0 gives -inf, x < 0 NaN, +inf +inf.
//   Ops: 2 frexp, 1 cmp/select pair, 2 add, 1 div, 1 mul, 13 fma.
#pragma once
#include "wave_target.h"

namespace tgr {

// Operation counts in the FLOP convention of the roofline (add, sub, mul,
// div, rint, ldexp, frexp = 1 flop; fma = 2 flops; compares and selects = 0).
constexpr double kFlopsExp = 1 + 1 + 15 * 2 + 1;       // 33
constexpr double kFlopsLog = 2 + 2 + 1 + 1 + 13 * 2;   // 32

__host__ __device__ __forceinline__ double fexp(double x) {
  const double kLog2e = 1.4426950408889634074;
  const double kLn2Hi = 6.93147180369123816490e-01;  // 32 significant bits
  const double kLn2Lo = 1.90821492927058770002e-10;
  // clamp (maxnum/minnum: a NaN becomes -746 here and is restored at the end)
  const double xc = fmin(fmax(x, -746.0), 710.0);
  const double k = rint(xc * kLog2e);
  double r = fma(-k, kLn2Hi, xc);
  r = fma(-k, kLn2Lo, r);
  double p = 1.6059043836821614599e-10;      // 1/13!
  p = fma(p, r, 2.0876756987868098979e-09);  // 1/12!
  p = fma(p, r, 2.5052108385441718775e-08);  // 1/11!
  p = fma(p, r, 2.7557319223985890653e-07);  // 1/10!
  p = fma(p, r, 2.7557319223985892511e-06);  // 1/9!
  p = fma(p, r, 2.4801587301587301566e-05);  // 1/8!
  p = fma(p, r, 1.9841269841269841253e-04);  // 1/7!
  p = fma(p, r, 1.3888888888888888889e-03);  // 1/6!
  p = fma(p, r, 8.3333333333333333333e-03);  // 1/5!
  p = fma(p, r, 4.1666666666666666667e-02);  // 1/4!
  p = fma(p, r, 1.6666666666666666667e-01);  // 1/3!
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  const double v = ldexp(p, (int)k);
  return x == x ? v : x;
}

__host__ __device__ __forceinline__ double flog(double x) {
  const double kLn2Hi = 6.93147180369123816490e-01;
  const double kLn2Lo = 1.90821492927058770002e-10;
  int e;
  double m = frexp(x, &e);  // m in [0.5, 1)
  const bool lo = m < 0.70710678118654752440;
  m = lo ? m + m : m;
  e = lo ? e - 1 : e;
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, z, 1.0 / 19.0);
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  p = fma(p, z, 1.0);
  const double de = (double)e;
  const double t = fma(s + s, p, de * kLn2Lo);
  double y = fma(de, kLn2Hi, t);
  // special values: frexp leaves 0, inf and NaN unchanged
  if (!(x > 0.0) || x == INFINITY) y = (x == 0.0) ? -INFINITY : (x == INFINITY ? INFINITY : NAN);
  return y;
}

}  // namespace tgr
