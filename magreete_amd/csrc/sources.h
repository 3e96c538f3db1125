// Incident fields of plane waves and Gaussian beams at one point, shared by
// the source kernel and the propagation's fused epilogue.
#pragma once
#include "mg_common.h"

struct mg_src_d2 { double x, y; };

// f(x) = exp(i k0 r_par) (plane, type 0) or the paraxial Gaussian beam
// exp(i k0 r_par - r_perp^2 / (w^2 (1 + i a))) / sqrt(1 + i a), a = 2 r_perp / (w^2 k0)
// (type 1).  2D: dir = (cos(-theta), sin(-theta)), r_par = c x - s y,
// r_perp = s x + c y (Transmission2D.py:118-122, 629-633); 3D: dir = u,
// r_par = x . u, r_perp = |x - r_par u| (Transmission3D.py:96-100).
MG_HD mg_src_d2 mg_source_value(int space, int type, const double* x, const double* dir, double k0, double w) {
  double rpar, rperp;
  if (space == 2) {
    const double c = dir[0], s = dir[1];
    rpar = c * x[0] - s * x[1];
    rperp = s * x[0] + c * x[1];
  } else {
    const double ux = dir[0], uy = dir[1], uz = dir[2];
    rpar = x[0] * ux + x[1] * uy + x[2] * uz;
    const double ex = x[0] - rpar * ux, ey = x[1] - rpar * uy, ez = x[2] - rpar * uz;
    rperp = sqrt(ex * ex + ey * ey + ez * ez);
  }
  if (type == 0) {
    double sn, cs;
    sincos(rpar * k0, &sn, &cs);
    return {cs, sn};
  }
  const double a = 2.0 * rperp / (w * w * k0);
  // q = rperp^2 / (w^2 (1 + i a)) = rperp^2/w^2 * (1 - i a)/(1 + a^2)
  const double f = rperp * rperp / (w * w) / (1.0 + a * a);
  const double zr = -f, zi = rpar * k0 + f * a;
  double sn, cs;
  sincos(zi, &sn, &cs);
  const double mag = exp(zr);
  // principal sqrt(1 + i a)
  const double r = sqrt(1.0 + a * a);
  const double sr = sqrt(0.5 * (r + 1.0));
  const double si = 0.5 * a / sr;   // sr * si = a / 2, no cancellation at small a
  const double nr = mag * cs, ni = mag * sn;
  const double den = sr * sr + si * si;
  return {(nr * sr + ni * si) / den, (ni * sr - nr * si) / den};
}
