// Pairwise Green's-function blocks for the four kinds, FP64, host and device.
// The 2D kinds evaluate Bessel functions as the reference's torch does
// (bessel_ref.h) unless MG_F_BESSEL_EXACT selects bessel.h.
//
// Each function takes the difference vector r = x_row - x_col and writes the
// d x d block g[a*d + b] (complex, as re/im pairs).  Coincident pairs give
// NaN + NaN i in every entry, the places where the reference's torch
// expressions give NaN (0/0 in R̂R̂, H0(0), exp(0)/0), which its fix-up code
// later detects with isnan.
#pragma once
#include "mg_common.h"
#include "bessel.h"
#include "bessel_ref.h"

struct mg_c { double re, im; };

MG_HD mg_c mg_cmul(mg_c a, mg_c b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

MG_HD double mg_wrap(double x) {
  // ((x + 0.5) mod 1) - 0.5 with a floor modulus, as torch's %= on floats
  const double y = x + 0.5;
  return (y - floor(y)) - 0.5;
}

// 2D TM: 0.25i H0(k0 R) (Transmission2D.py:595-618); regularize compares R
// (unscaled) with radius.
MG_HD void mg_green_2d_tm(double rx, double ry, double k0, int flags, double radius, mg_c* g) {
  if (flags & MG_F_PERIODIC_Y) ry = mg_wrap(ry);
  else if (flags & MG_F_PERIODIC_X) rx = mg_wrap(rx);
  double R = sqrt(rx * rx + ry * ry);
  if ((flags & MG_F_REGULARIZE) && R < radius) R = 0.0;
  if (R == 0.0) { g[0] = {NAN, NAN}; return; }
  double j0, y0;
  if (flags & MG_F_BESSEL_EXACT) {
    mg_bessel0(R * k0, j0, y0);
  } else {
    j0 = mg_ref_j0(R * k0);
    y0 = mg_ref_y0(R * k0);
  }
  g[0] = {-0.25 * y0, 0.25 * j0};
}

// 2D TE dyadic 0.25i[(I - R̂R̂) H0(kR) - (I - 2 R̂R̂) H1(kR)/(kR)]
// (Transmission2D.py:63-91): R̂R̂ is formed from the unscaled R, and the
// regularisation compares the scaled k0 R with radius.
MG_HD void mg_green_2d_te(double rx, double ry, double k0, int flags, double radius, mg_c* g) {
  if (flags & MG_F_PERIODIC_Y) ry = mg_wrap(ry);
  else if (flags & MG_F_PERIODIC_X) rx = mg_wrap(rx);
  const double R = sqrt(rx * rx + ry * ry);
  double kr = R * k0;
  if ((flags & MG_F_REGULARIZE) && kr < radius) kr = 0.0;
  if (R == 0.0 || kr == 0.0) {
    for (int e = 0; e < 4; ++e) g[e] = {NAN, NAN};
    return;
  }
  const double inv = 1.0 / (R * R);
  const double xx = rx * rx * inv, xy = rx * ry * inv, yy = ry * ry * inv;
  double j0, y0, j1, y1;
  if (flags & MG_F_BESSEL_EXACT) {
    mg_bessel01(kr, j0, y0, j1, y1);
  } else {
    j0 = mg_ref_j0(kr); y0 = mg_ref_y0(kr);
    j1 = mg_ref_j1(kr); y1 = mg_ref_y1(kr);
  }
  // H0 = j0 + i y0 ; H1/kr
  const mg_c h0 = {j0, y0};
  const mg_c h1 = {j1 / kr, y1 / kr};
  // entry = 0.25 i [(d_ab - p_ab) h0 - (d_ab - 2 p_ab) h1]
  const double p[4] = {xx, xy, xy, yy};
  for (int e = 0; e < 4; ++e) {
    const double dl = (e == 0 || e == 3) ? 1.0 : 0.0;
    const double a = dl - p[e], c = dl - 2.0 * p[e];
    const double re = a * h0.re - c * h1.re;
    const double im = a * h0.im - c * h1.im;
    g[e] = {-0.25 * im, 0.25 * re};
  }
}

// 3D dyadic [(I - R̂R̂) - (I - 3 R̂R̂)(1/(ikR) + 1/(kR)^2)] e^{ikR}/(4 pi R)
// (Transmission3D.py:47-77).
MG_HD void mg_green_3d_vector(double rx, double ry, double rz, double k0, int flags, double radius, mg_c* g) {
  if (flags & MG_F_PERIODIC_X) rx = mg_wrap(rx);
  if (flags & MG_F_PERIODIC_Y) ry = mg_wrap(ry);
  if (flags & MG_F_PERIODIC_Z) rz = mg_wrap(rz);
  double R = sqrt(rx * rx + ry * ry + rz * rz);
  const bool sing = (R == 0.0) || ((flags & MG_F_REGULARIZE) && R < radius);
  if (sing) {
    for (int e = 0; e < 9; ++e) g[e] = {NAN, NAN};
    return;
  }
  const double inv = 1.0 / (R * R);
  const double v[3] = {rx, ry, rz};
  const double kr = k0 * R;
  const double ikr = 1.0 / kr;
  // f = 1/(i kR) + 1/(kR)^2 = ikr^2 - i ikr
  const double fre = ikr * ikr, fim = -ikr;
  double sn, cs;
  sincos(kr, &sn, &cs);
  const double w = 1.0 / (4.0 * 3.14159265358979323846 * R);
  const mg_c ph = {cs * w, sn * w};
  for (int a = 0; a < 3; ++a) {
    for (int b = 0; b < 3; ++b) {
      const double p = v[a] * v[b] * inv;
      const double dl = (a == b) ? 1.0 : 0.0;
      const double A = dl - p, C = dl - 3.0 * p;
      const mg_c blk = {A - C * fre, -C * fim};
      g[a * 3 + b] = mg_cmul(blk, ph);
    }
  }
}

// 3D scalar e^{ikR}/(4 pi R) (Transmission3D.py:573-601).
MG_HD void mg_green_3d_scalar(double rx, double ry, double rz, double k0, int flags, double radius, mg_c* g) {
  if (flags & MG_F_PERIODIC_X) rx = mg_wrap(rx);
  if (flags & MG_F_PERIODIC_Y) ry = mg_wrap(ry);
  if (flags & MG_F_PERIODIC_Z) rz = mg_wrap(rz);
  double R = sqrt(rx * rx + ry * ry + rz * rz);
  if (R == 0.0 || ((flags & MG_F_REGULARIZE) && R < radius)) { g[0] = {NAN, NAN}; return; }
  double sn, cs;
  sincos(k0 * R, &sn, &cs);
  const double w = 1.0 / (4.0 * 3.14159265358979323846 * R);
  g[0] = {cs * w, sn * w};
}

// Dispatch on kind; xr/xc point at the row/column point coordinates.
template <int KIND>
MG_HD void mg_green(const double* xr, const double* xc, double k0, int flags, double radius, mg_c* g) {
  if (KIND == MG_2D_TM) mg_green_2d_tm(xr[0] - xc[0], xr[1] - xc[1], k0, flags, radius, g);
  else if (KIND == MG_2D_TE) mg_green_2d_te(xr[0] - xc[0], xr[1] - xc[1], k0, flags, radius, g);
  else if (KIND == MG_3D_VECTOR) mg_green_3d_vector(xr[0] - xc[0], xr[1] - xc[1], xr[2] - xc[2], k0, flags, radius, g);
  else mg_green_3d_scalar(xr[0] - xc[0], xr[1] - xc[1], xr[2] - xc[2], k0, flags, radius, g);
}
