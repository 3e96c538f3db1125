// FP64 Bessel functions J0, Y0, J1, Y1 for x >= 0, host and device.
//
// Written from the series/asymptotic structure of Abramowitz & Stegun:
//  * x <= 8: J0, J1/x and the remainders Y0 - (2/pi) ln(x/2) J0 and
//    (Y1 - (2/pi)(ln(x/2) J1 - 1/x))/x are entire in x^2 (A&S 9.1.10-9.1.13),
//    so each is a Chebyshev series in s = x^2/32 - 1.  The ln(x/2) term of
//    Y0 is explicit, which keeps the small-argument log exact.
//  * x > 8: Hankel's modulus/phase form (A&S 9.2.5-9.2.6) with P_n and
//    Q_n / (8/x) as Chebyshev series in s = 2 (8/x)^2 - 1; the phase
//    x - (2n+1) pi/4 is applied by rotating sincos(x), with no reduction of
//    x - pi/4.
// The coefficients are this project's own mpmath interpolants
// (gen_bessel.py -> bessel_coeffs.h).
#pragma once
#include "mg_common.h"
#include "bessel_coeffs.h"

template <int N>
MG_HD double mg_cheb(const double (&c)[N], double s) {
  // Clenshaw recurrence b_k = c_k + 2 s b_{k+1} - b_{k+2}
  double b1 = 0.0, b2 = 0.0;
  const double s2 = s + s;
#pragma unroll
  for (int k = N - 1; k >= 1; --k) {
    const double b0 = fma(s2, b1, c[k] - b2);
    b2 = b1;
    b1 = b0;
  }
  return fma(s, b1, c[0] - b2);
}

// J0, Y0, J1, Y1 at x >= 0.  At x == 0: J0 = 1, J1 = 0, Y0 = Y1 = -inf.
MG_HD void mg_bessel01(double x, double& j0, double& y0, double& j1, double& y1) {
  constexpr double kTwoOverPi = 0.63661977236758134308;
  constexpr double kInvSqrt2 = 0.70710678118654752440;
  if (x == 0.0) {
    j0 = 1.0; j1 = 0.0;
    y0 = -INFINITY; y1 = -INFINITY;
    return;
  }
  if (x <= 8.0) {
    const double s = x * x * (1.0 / 32.0) - 1.0;
    j0 = mg_cheb(BJ0, s);
    const double lg = kTwoOverPi * log(0.5 * x);
    y0 = fma(lg, j0, mg_cheb(BR0, s));
    j1 = x * mg_cheb(BJ1, s);
    y1 = fma(lg, j1, fma(x, mg_cheb(BR1, s), -kTwoOverPi / x));
  } else {
    const double z = 8.0 / x;
    const double s = 2.0 * z * z - 1.0;
    const double p0 = mg_cheb(BP0, s), q0 = z * mg_cheb(BQ0, s);
    const double p1 = mg_cheb(BP1, s), q1 = z * mg_cheb(BQ1, s);
    double sn, cs;
    sincos(x, &sn, &cs);
    const double a = sqrt(kTwoOverPi / x);
    // chi0 = x - pi/4, chi1 = x - 3 pi/4
    const double c0 = (cs + sn) * kInvSqrt2, s0 = (sn - cs) * kInvSqrt2;
    const double c1 = s0, s1 = -c0;
    j0 = a * (p0 * c0 - q0 * s0);
    y0 = a * (p0 * s0 + q0 * c0);
    j1 = a * (p1 * c1 - q1 * s1);
    y1 = a * (p1 * s1 + q1 * c1);
  }
}

// J0 and Y0 only (2D TM needs H0 alone).
MG_HD void mg_bessel0(double x, double& j0, double& y0) {
  constexpr double kTwoOverPi = 0.63661977236758134308;
  constexpr double kInvSqrt2 = 0.70710678118654752440;
  if (x == 0.0) {
    j0 = 1.0; y0 = -INFINITY;
    return;
  }
  if (x <= 8.0) {
    const double s = x * x * (1.0 / 32.0) - 1.0;
    j0 = mg_cheb(BJ0, s);
    y0 = fma(kTwoOverPi * log(0.5 * x), j0, mg_cheb(BR0, s));
  } else {
    const double z = 8.0 / x;
    const double s = 2.0 * z * z - 1.0;
    const double p0 = mg_cheb(BP0, s), q0 = z * mg_cheb(BQ0, s);
    double sn, cs;
    sincos(x, &sn, &cs);
    const double a = sqrt(kTwoOverPi / x);
    const double c0 = (cs + sn) * kInvSqrt2, s0 = (sn - cs) * kInvSqrt2;
    j0 = a * (p0 * c0 - q0 * s0);
    y0 = a * (p0 * s0 + q0 * c0);
  }
}
