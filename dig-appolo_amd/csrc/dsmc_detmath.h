// Deterministic elementary functions shared by the gfx950 kernels and host
// code.  Built only from IEEE-754 +, -, *, /, floor and bit manipulation, with
// floating-point contraction disabled, so host and device return the same
// bits for the same argument.  That is what lets the tracking kernel (wall
// reflection, inflow sampling) and the collision kernel reproduce the host
// restatement exactly.
//
// Methods (all published, coefficients are the exact Taylor coefficients of
// the named series, folded by the compiler):
//   det_log   : reduction x = 2^k m, m in [sqrt(1/2), sqrt(2)); log m via the
//               atanh series in s = f/(2+f), assembled as in Sun's fdlibm
//               e_log.c (log(1+f) = f - (hfsq - s(hfsq+R))), ln2 split into
//               hi/lo (Cody & Waite, "Software Manual for the Elementary
//               Functions", 1980).
//   det_exp   : Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, Taylor
//               series of e^r to degree 14, exact scaling by 2^k.
//   det_sincos2pi : sin/cos of 2*pi*u, exact reduction of u to the nearest
//               octant (u*8 - round(u*8) is exact), Taylor series on
//               |a| <= pi/8, octant rotation with cos(pi/4) = sqrt(1/2).
//   det_pow   : exp(y log x) for x > 0.
// Accuracy is a few ulp (checked against glibc in the tests); what matters
// here is that the results are identical on host and device.
#pragma once
#include <stdint.h>
#include <string.h>
#include <math.h>
#include "dsmc_rng.h"

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

namespace dsmc {

DSMC_HD uint64_t dbits(double x) {
  uint64_t u;
  memcpy(&u, &x, sizeof(u));
  return u;
}
DSMC_HD double bitsd(uint64_t u) {
  double x;
  memcpy(&x, &u, sizeof(x));
  return x;
}

// 2^k * x, k in [-2044, 2046]; exact unless the result is subnormal/overflows.
DSMC_HD double det_scalbn(double x, int k) {
  if (k > 1023) {
    x *= bitsd((uint64_t)(1023 + 1023) << 52);
    k -= 1023;
    if (k > 1023) k = 1023;
  } else if (k < -1022) {
    x *= bitsd((uint64_t)(-1022 + 54 + 1023) << 52);  // 2^-968
    k += 968;
    if (k < -1022) k = -1022;
  }
  return x * bitsd((uint64_t)(k + 1023) << 52);
}

DSMC_HD double det_log(double x) {
  const double LN2_HI = 6.93147180369123816490e-01;  // ln2 upper 32 bits
  const double LN2_LO = 1.90821492927058770002e-10;  // ln2 - LN2_HI
  const double SQRT2 = 1.41421356237309514547;
  if (!(x > 0.0)) {
    if (x == 0.0) return -HUGE_VAL;
    return (x - x) / 0.0;  // NaN for negative or NaN input
  }
  if (x == HUGE_VAL) return x;
  int k = 0;
  uint64_t b = dbits(x);
  if ((b >> 52) == 0) {  // subnormal: scale into the normal range
    x *= 18014398509481984.0;  // 2^54
    b = dbits(x);
    k = -54;
  }
  k += (int)((b >> 52) & 0x7ff) - 1023;
  double m = bitsd((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  // [1,2)
  if (m > SQRT2) {
    m = m * 0.5;
    k += 1;
  }
  const double f = m - 1.0;  // exact
  const double s = f / (2.0 + f);
  const double z = s * s;
  // R = sum_{n>=1} 2 z^n / (2n+1), truncated at n = 11 (|z| <= 0.0295).
  const double R =
      z * ((2.0 / 3.0) +
      z * ((2.0 / 5.0) +
      z * ((2.0 / 7.0) +
      z * ((2.0 / 9.0) +
      z * ((2.0 / 11.0) +
      z * ((2.0 / 13.0) +
      z * ((2.0 / 15.0) +
      z * ((2.0 / 17.0) +
      z * ((2.0 / 19.0) +
      z * ((2.0 / 21.0) +
      z * (2.0 / 23.0)))))))))));
  const double hfsq = 0.5 * f * f;
  const double dk = (double)k;
  return dk * LN2_HI - ((hfsq - (s * (hfsq + R) + dk * LN2_LO)) - f);
}

DSMC_HD double det_exp(double x) {
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  const double INV_LN2 = 1.44269504088896338700e+00;
  if (x != x) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -745.1332191019412) return 0.0;
  const double kd = floor(x * INV_LN2 + 0.5);
  const double hi = x - kd * LN2_HI;  // kd*LN2_HI is exact (|kd| < 2^11)
  const double lo = kd * LN2_LO;
  const double r = hi - lo;
  const double p =
      1.0 + r * (1.0 +
      r * ((1.0 / 2.0) +
      r * ((1.0 / 6.0) +
      r * ((1.0 / 24.0) +
      r * ((1.0 / 120.0) +
      r * ((1.0 / 720.0) +
      r * ((1.0 / 5040.0) +
      r * ((1.0 / 40320.0) +
      r * ((1.0 / 362880.0) +
      r * ((1.0 / 3628800.0) +
      r * ((1.0 / 39916800.0) +
      r * ((1.0 / 479001600.0) +
      r * ((1.0 / 6227020800.0) +
      r * (1.0 / 87178291200.0))))))))))))));
  return det_scalbn(p, (int)kd);
}

// s = sin(2 pi u), c = cos(2 pi u) for finite |u| < 2^49.
DSMC_HD void det_sincos2pi(double u, double& s, double& c) {
  const double PI_4 = 7.85398163397448278999e-01;
  const double SQRT1_2 = 7.07106781186547572737e-01;
  const double t = u * 8.0;            // exact
  const double q = floor(t + 0.5);
  const double r8 = t - q;             // exact, in [-1/2, 1/2]
  const long long iql = (long long)q;
  const int iq = (int)(iql & 7);
  const double a = r8 * PI_4;          // |a| <= pi/8
  const double a2 = a * a;
  const double sa = a + a * a2 * ((-1.0 / 6.0) +
      a2 * ((1.0 / 120.0) +
      a2 * ((-1.0 / 5040.0) +
      a2 * ((1.0 / 362880.0) +
      a2 * ((-1.0 / 39916800.0) +
      a2 * ((1.0 / 6227020800.0) +
      a2 * (-1.0 / 1307674368000.0)))))));
  const double ca = 1.0 + a2 * ((-1.0 / 2.0) +
      a2 * ((1.0 / 24.0) +
      a2 * ((-1.0 / 720.0) +
      a2 * ((1.0 / 40320.0) +
      a2 * ((-1.0 / 3628800.0) +
      a2 * ((1.0 / 479001600.0) +
      a2 * ((-1.0 / 87178291200.0) +
      a2 * (1.0 / 20922789888000.0))))))));
  switch (iq) {
    case 0: s = sa; c = ca; break;
    case 1: s = SQRT1_2 * (ca + sa); c = SQRT1_2 * (ca - sa); break;
    case 2: s = ca; c = -sa; break;
    case 3: s = SQRT1_2 * (ca - sa); c = -(SQRT1_2 * (ca + sa)); break;
    case 4: s = -sa; c = -ca; break;
    case 5: s = -(SQRT1_2 * (ca + sa)); c = -(SQRT1_2 * (ca - sa)); break;
    case 6: s = -ca; c = sa; break;
    default: s = -(SQRT1_2 * (ca - sa)); c = SQRT1_2 * (ca + sa); break;
  }
}

// x^y for x > 0 (the only use in the engine).
DSMC_HD double det_pow(double x, double y) {
  if (y == 0.0 || x == 1.0) return 1.0;
  if (y == 1.0) return x;
  return det_exp(y * det_log(x));
}

}  // namespace dsmc
