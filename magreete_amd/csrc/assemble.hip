// Green's-tensor assembly straight into a complex matrix.
//
// out[(d i + a) ld + d j + b] = scale * G_ab(x_row[i] - x_col[j]), with the
// fix-ups the reference applies after building G:
//  * MG_F_ZERO_SELF: block (i, i) is zero (G0(None) zeroes self blocks,
//    Transmission3D.py:261-264, Transmission2D.py:276-278);
//  * MG_F_DIAG: entry (r, r) is `diag` (fill_diagonal_(1) followed by the
//    self-interaction shift, Transmission3D.py:197-203);
//  * MG_F_NANFILL: NaN entries become `fill` (LDOS measurement fix-up,
//    Transmission3D.py:353-362).
// One thread per (i, j) pair writes the d x d block; consecutive threads take
// consecutive j, so each block row is one coalesced run.  The FP64 matrix
// and an optional complex-float copy are written in the same pass.
#include "green.h"
#include "mg_internal.h"

namespace {

template <int KIND>
__global__ __launch_bounds__(256) void k_assemble(const double* __restrict__ rows, int64_t nrows,
                                                  const double* __restrict__ cols, int64_t ncols, double k0,
                                                  double2 scale, double2 diag, double2 fill, int flags,
                                                  double radius, double2* __restrict__ out, int64_t ld,
                                                  float2* __restrict__ out32, int64_t ld32) {
  constexpr int D = (KIND == MG_2D_TE) ? 2 : (KIND == MG_3D_VECTOR ? 3 : 1);
  constexpr int SP = (KIND == MG_2D_TM || KIND == MG_2D_TE) ? 2 : 3;
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncols) return;
  double xc[3];
#pragma unroll
  for (int t = 0; t < SP; ++t) xc[t] = cols[j * SP + t];
  for (int64_t i = blockIdx.y; i < nrows; i += gridDim.y) {
    double xr[3];
#pragma unroll
    for (int t = 0; t < SP; ++t) xr[t] = rows[i * SP + t];
    mg_c g[D * D];
    const bool self = i == j;
    if ((flags & MG_F_ZERO_SELF) && self) {
#pragma unroll
      for (int t = 0; t < D * D; ++t) g[t] = {0.0, 0.0};
    } else {
      mg_green<KIND>(xr, xc, k0, flags, radius, g);
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
#pragma unroll
      for (int b = 0; b < D; ++b) {
        const mg_c v = g[a * D + b];
        double2 o;
        o.x = scale.x * v.re - scale.y * v.im;
        o.y = scale.x * v.im + scale.y * v.re;
        if ((flags & MG_F_NANFILL) && (isnan(o.x) || isnan(o.y))) o = fill;
        if ((flags & MG_F_DIAG) && self && a == b) o = diag;
        const int64_t r = D * i + a, c = D * j + b;
        if (out) out[r * ld + c] = o;
        if (out32) out32[r * ld32 + c] = make_float2((float)o.x, (float)o.y);
      }
    }
  }
}

template <int KIND>
int launch_assemble(const double* rows, int64_t nrows, const double* cols, int64_t ncols, double k0, double2 scale,
                    double2 diag, double2 fill, int flags, double radius, double2* out, int64_t ld, float2* out32,
                    int64_t ld32, hipStream_t s) {
  const int64_t gx = (ncols + 255) / 256;
  const int64_t gy = nrows < 65535 ? nrows : 65535;
  hipLaunchKernelGGL(k_assemble<KIND>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, rows, nrows, cols, ncols,
                     k0, scale, diag, fill, flags, radius, out, ld, out32, ld32);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

}  // namespace

extern "C" int mg_assemble(int kind, const double* rows, int64_t nrows, const double* cols, int64_t ncols, double k0,
                           double scale_re, double scale_im, double diag_re, double diag_im, double fill_re,
                           double fill_im, int flags, double radius, void* out_c128, int64_t ld_c128,
                           void* out_c64, int64_t ld_c64, void* stream) {
  if (nrows < 0 || ncols < 0 || (!out_c128 && !out_c64)) return MG_ERR_ARG;
  if (nrows == 0 || ncols == 0) return MG_OK;
  const int d = mg_kind_dim(kind);
  if (out_c128 && ld_c128 < d * ncols) return MG_ERR_ARG;
  if (out_c64 && ld_c64 < d * ncols) return MG_ERR_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const double2 sc = make_double2(scale_re, scale_im), dg = make_double2(diag_re, diag_im),
                fl = make_double2(fill_re, fill_im);
  double2* o = reinterpret_cast<double2*>(out_c128);
  float2* o32 = reinterpret_cast<float2*>(out_c64);
  switch (kind) {
    case MG_2D_TM: return launch_assemble<MG_2D_TM>(rows, nrows, cols, ncols, k0, sc, dg, fl, flags, radius, o, ld_c128, o32, ld_c64, s);
    case MG_2D_TE: return launch_assemble<MG_2D_TE>(rows, nrows, cols, ncols, k0, sc, dg, fl, flags, radius, o, ld_c128, o32, ld_c64, s);
    case MG_3D_VECTOR: return launch_assemble<MG_3D_VECTOR>(rows, nrows, cols, ncols, k0, sc, dg, fl, flags, radius, o, ld_c128, o32, ld_c64, s);
    case MG_3D_SCALAR: return launch_assemble<MG_3D_SCALAR>(rows, nrows, cols, ncols, k0, sc, dg, fl, flags, radius, o, ld_c128, o32, ld_c64, s);
    default: return MG_ERR_KIND;
  }
}
