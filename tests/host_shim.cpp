// Host build of the kernels' special-function and Green's-function code, so
// the CPU tests check the same source the GPU kernels compile.
#include "bessel.h"
#include "green.h"

extern "C" void shim_bessel01(const double* x, int n, double* out) {
  for (int i = 0; i < n; ++i) {
    double a, b, c, d;
    mg_bessel01(x[i], a, b, c, d);
    out[4 * i] = a;
    out[4 * i + 1] = b;
    out[4 * i + 2] = c;
    out[4 * i + 3] = d;
  }
}

extern "C" void shim_bessel_ref(const double* x, int n, double* out) {
  for (int i = 0; i < n; ++i) {
    out[4 * i] = mg_ref_j0(x[i]);
    out[4 * i + 1] = mg_ref_y0(x[i]);
    out[4 * i + 2] = mg_ref_j1(x[i]);
    out[4 * i + 3] = mg_ref_y1(x[i]);
  }
}

// g: d*d complex (re, im) for the difference vector r.
extern "C" void shim_green(int kind, const double* r, double k0, int flags, double radius, double* g) {
  const double zero[3] = {0.0, 0.0, 0.0};
  mg_c out[9];
  switch (kind) {
    case MG_2D_TM: mg_green<MG_2D_TM>(r, zero, k0, flags, radius, out); break;
    case MG_2D_TE: mg_green<MG_2D_TE>(r, zero, k0, flags, radius, out); break;
    case MG_3D_VECTOR: mg_green<MG_3D_VECTOR>(r, zero, k0, flags, radius, out); break;
    default: mg_green<MG_3D_SCALAR>(r, zero, k0, flags, radius, out); break;
  }
  const int d = mg_kind_dim(kind);
  for (int e = 0; e < d * d; ++e) {
    g[2 * e] = out[e].re;
    g[2 * e + 1] = out[e].im;
  }
}
