// J0, J1, Y0, Y1 evaluated exactly as the reference's numerics produce them.
//
// The reference takes its Bessel functions from torch.special
// (Transmission2D.py:56-59, 588-591), i.e. torch's restatement of the Cephes
// rational approximations (S. L. Moshier).  That restatement evaluates every
// polynomial by Horner over its stored coefficients, so the denominators that
// Cephes evaluates with an implicit leading 1 (RQ, QQ, YQ) lose that term.
// The functions are therefore off by up to ~2e-6 relative for x in about
// [3.4, 12] (checked against mpmath in tests/test_special_functions.py).
// The drop-in reproduces them by default so that its fields match the
// reference; MG_F_BESSEL_EXACT selects the accurate evaluation of bessel.h.
//
// Branches: x <= 5, rational functions of x^2 with the first two squared
// zeros of J_n factored out (J) or a log term added (Y); x > 5, Hankel's
// P/Q form in z = 25/x^2 with phase x - (2n+1) pi/4.
#pragma once
#include "mg_common.h"
#include "bessel_ref_coeffs.h"

template <int N>
MG_HD double mg_horner(const double (&c)[N], double z) {
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) r = r * z + c[i];
  return r;
}

constexpr double kRefPiOver4 = 0.785398163397448309615660845819875721;
constexpr double kRef3PiOver4 = 2.356194490192344928846982537459627163;
constexpr double kRefSqrt2OverPi = 0.797884560802865355879892119868763737;
constexpr double kRefTwoOverPi = 0.636619772367581343075535053490057448;

MG_HD double mg_ref_j0(double x) {
  if (x < 0) x = -x;
  if (x <= 5.0) {
    if (x < 0.00001) return 1.0 - x * x / 4.0;
    const double z = x * x;
    const double rp = mg_horner(REF_J0_RP, z), rq = mg_horner(REF_J0_RQ, z);
    return (z - REF_J0_ZERO1_SQ) * (z - REF_J0_ZERO2_SQ) * rp / rq;
  }
  const double z = 25.0 / (x * x);
  const double pp = mg_horner(REF_J0_PP, z), pq = mg_horner(REF_J0_PQ, z);
  const double qp = mg_horner(REF_J0_QP, z), qq = mg_horner(REF_J0_QQ, z);
  return (pp / pq * cos(x - kRefPiOver4) - 5.0 / x * (qp / qq) * sin(x - kRefPiOver4)) * kRefSqrt2OverPi / sqrt(x);
}

MG_HD double mg_ref_j1_pos(double x) {
  if (x <= 5.0) {
    const double z = x * x;
    const double rp = mg_horner(REF_J1_RP, z), rq = mg_horner(REF_J1_RQ, z);
    return rp / rq * x * (z - REF_J1_ZERO1_SQ) * (z - REF_J1_ZERO2_SQ);
  }
  const double z = 5.0 / x * (5.0 / x);
  const double pp = mg_horner(REF_J1_PP, z), pq = mg_horner(REF_J1_PQ, z);
  const double qp = mg_horner(REF_J1_QP, z), qq = mg_horner(REF_J1_QQ, z);
  return (pp / pq * cos(x - kRef3PiOver4) - 5.0 / x * (qp / qq) * sin(x - kRef3PiOver4)) * kRefSqrt2OverPi / sqrt(x);
}

MG_HD double mg_ref_j1(double x) { return x < 0 ? -mg_ref_j1_pos(-x) : mg_ref_j1_pos(x); }

MG_HD double mg_ref_y0(double x) {
  if (x <= 5.0) {
    if (x == 0.0) return -INFINITY;
    if (x < 0.0) return NAN;
    const double z = x * x;
    const double yp = mg_horner(REF_Y0_YP, z), yq = mg_horner(REF_Y0_YQ, z);
    return yp / yq + (kRefTwoOverPi * log(x) * mg_ref_j0(x));
  }
  const double z = 25.0 / (x * x);
  const double pp = mg_horner(REF_Y0_PP, z), pq = mg_horner(REF_Y0_PQ, z);
  const double qp = mg_horner(REF_Y0_QP, z), qq = mg_horner(REF_Y0_QQ, z);
  return (pp / pq * sin(x - kRefPiOver4) + 5.0 / x * (qp / qq) * cos(x - kRefPiOver4)) * kRefSqrt2OverPi / sqrt(x);
}

MG_HD double mg_ref_y1(double x) {
  if (x <= 5.0) {
    if (x == 0.0) return -INFINITY;
    if (x <= 0.0) return NAN;
    const double z = x * x;
    const double yp = mg_horner(REF_Y1_YP, z), yq = mg_horner(REF_Y1_YQ, z);
    return x * (yp / yq) + (kRefTwoOverPi * (mg_ref_j1(x) * log(x) - 1.0 / x));
  }
  const double z = 5.0 / x * (5.0 / x);
  const double pp = mg_horner(REF_Y1_PP, z), pq = mg_horner(REF_Y1_PQ, z);
  const double qp = mg_horner(REF_Y1_QP, z), qq = mg_horner(REF_Y1_QQ, z);
  return (pp / pq * sin(x - kRef3PiOver4) + 5.0 / x * (qp / qq) * cos(x - kRef3PiOver4)) * kRefSqrt2OverPi / sqrt(x);
}
