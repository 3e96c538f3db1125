// Small field kernels: Hankel functions, incident plane waves and Gaussian
// beams, and column dot products for the LDOS.
#include "green.h"
#include "mg_internal.h"
#include "sources.h"

namespace {

__global__ void k_hankel1(int nu, int exact, const double* __restrict__ x, int64_t n, double2* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = x[i];
  if (exact) {
    double j0, y0, j1, y1;
    mg_bessel01(v, j0, y0, j1, y1);
    out[i] = nu == 0 ? make_double2(j0, y0) : make_double2(j1, y1);
  } else {
    out[i] = nu == 0 ? make_double2(mg_ref_j0(v), mg_ref_y0(v)) : make_double2(mg_ref_j1(v), mg_ref_y1(v));
  }
}

// E[m, a, t] = f_t(x_m) * pol[t, a] with f the plane wave or Gaussian beam
// of sources.h (dirs[t]: 2D (cos(-theta), sin(-theta)), 3D the direction).
__global__ void k_sources(int space, int type, const double* __restrict__ pts, int64_t npts,
                          const double* __restrict__ dirs, int64_t T, const double* __restrict__ pol, int d,
                          double k0, double w, double2* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  for (int64_t m = blockIdx.y; m < npts; m += gridDim.y) {
    const mg_src_d2 v = mg_source_value(space, type, pts + space * m, dirs + space * t, k0, w);
    const double2 e = make_double2(v.x, v.y);
    if (d == 1 && pol == nullptr) {
      out[m * T + t] = e;
    } else {
      for (int a = 0; a < d; ++a) {
        const double pa = pol[t * d + a];
        out[(m * d + a) * T + t] = make_double2(e.x * pa, e.y * pa);
      }
    }
  }
}

// out[c] = sum_r Y[r, c] X[r, c] (no conjugation), one workgroup per 64
// columns, rows split over 4 waves and reduced in LDS.
__global__ __launch_bounds__(256) void k_coldot(const double2* __restrict__ Y, int64_t ldy,
                                                const double2* __restrict__ X, int64_t ldx, int64_t n,
                                                int64_t ncols, double2* __restrict__ out) {
  __shared__ double2 part[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t c = (int64_t)blockIdx.x * 64 + lane;
  double sr = 0.0, si = 0.0;
  if (c < ncols) {
    for (int64_t r = wave; r < n; r += 4) {
      const double2 y = Y[r * ldy + c], x = X[r * ldx + c];
      sr += y.x * x.x - y.y * x.y;
      si += y.x * x.y + y.y * x.x;
    }
  }
  part[wave][lane] = make_double2(sr, si);
  __syncthreads();
  if (wave == 0 && c < ncols) {
    double2 acc = part[0][lane];
    for (int k = 1; k < 4; ++k) { acc.x += part[k][lane].x; acc.y += part[k][lane].y; }
    out[c] = acc;
  }
}

}  // namespace

extern "C" {

int mg_hankel1(int nu, int exact, const double* x, int64_t n, void* out, void* stream) {
  if ((nu != 0 && nu != 1) || n < 0) return MG_ERR_ARG;
  if (n == 0) return MG_OK;
  hipLaunchKernelGGL(k_hankel1, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), nu, exact, x, n, reinterpret_cast<double2*>(out));
  MG_HIP(hipGetLastError());
  return MG_OK;
}

int mg_coldot(const void* Y, int64_t ldy, const void* X, int64_t ldx, int64_t n, int64_t ncols, void* out,
              void* stream) {
  if (n < 0 || ncols < 0) return MG_ERR_ARG;
  if (ncols == 0) return MG_OK;
  hipLaunchKernelGGL(k_coldot, dim3((unsigned)((ncols + 63) / 64)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), reinterpret_cast<const double2*>(Y), ldy,
                     reinterpret_cast<const double2*>(X), ldx, n, ncols, reinterpret_cast<double2*>(out));
  MG_HIP(hipGetLastError());
  return MG_OK;
}

int mg_sources(int space, int type, const double* points, int64_t npts, const double* dirs, int64_t T,
               const double* pol, int64_t d, double k0, double w, void* out, void* stream) {
  if ((space != 2 && space != 3) || (type != 0 && type != 1) || npts < 0 || T < 0 || d < 1 || d > 3)
    return MG_ERR_ARG;
  if (d > 1 && !pol) return MG_ERR_ARG;
  if (npts == 0 || T == 0) return MG_OK;
  const dim3 grid((unsigned)((T + 63) / 64), (unsigned)(npts < 65535 ? npts : 65535));
  hipLaunchKernelGGL(k_sources, grid, dim3(64), 0, reinterpret_cast<hipStream_t>(stream), space, type, points, npts,
                     dirs, T, pol, (int)d, k0, w, reinterpret_cast<double2*>(out));
  MG_HIP(hipGetLastError());
  return MG_OK;
}

}  // extern "C"
