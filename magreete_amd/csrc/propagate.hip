// Fused propagation E = alpha * G(meas, scat) @ Ek + beta * E.
//
// Restates the matmul of propagate / propagate_ss (Transmission3D.py:163,
// 220; Transmission2D.py:190, 701): the (d M) x (d N) Green matrix is
// evaluated tile by tile in FP64 into LDS (GreenA loader) and contracted on
// FP64 MFMA against Ek; it never reaches HBM.  Coincident pairs evaluate to
// NaN, so the affected output rows are NaN exactly where the reference's are,
// and the caller applies the reference's row fix-up.
#include "cgemm.h"
#include "mg_internal.h"

namespace {

template <int KIND>
struct PropCfg {
  static constexpr int BM = 64, BN = 64, BK = 8, WM = 2, WN = 2;
};
template <>
struct PropCfg<MG_3D_VECTOR> {
  static constexpr int BM = 96, BN = 64, BK = 12, WM = 2, WN = 2;
};

template <int KIND>
__global__ __launch_bounds__(256) void k_propagate(int64_t M, int64_t N, int64_t K, double2 alpha, GreenA<KIND> g,
                                                   const void* Ek, int64_t ldek, double2 beta, void* E,
                                                   int64_t lde, int nbm, int nbn) {
  typedef PropCfg<KIND> P;
  int tm, tn;
  mg_tile_order(blockIdx.x, nbm, nbn, tm, tn);
  cgemm_body<double, P::BM, P::BN, P::BK, P::WM, P::WN, GreenA<KIND>>(M, N, K, alpha, g, Ek, ldek, beta, E, lde, tm,
                                                                      tn);
}

template <int KIND>
int launch(const double* meas, int64_t n_meas, const double* scat, int64_t n_scat, double k0, double2 alpha,
           int flags, double radius, const void* Ek, int64_t ldek, int64_t nrhs, double2 beta, void* E, int64_t lde,
           hipStream_t s) {
  typedef PropCfg<KIND> P;
  constexpr int D = (KIND == MG_2D_TE) ? 2 : (KIND == MG_3D_VECTOR ? 3 : 1);
  const int64_t M = D * n_meas, K = D * n_scat, N = nrhs;
  const int nbm = (int)((M + P::BM - 1) / P::BM), nbn = (int)((N + P::BN - 1) / P::BN);
  GreenA<KIND> g{meas, scat, n_meas, n_scat, k0, flags, radius};
  hipLaunchKernelGGL(k_propagate<KIND>, dim3(nbm * nbn), dim3(256), 0, s, M, N, K, alpha, g, Ek, ldek, beta, E, lde,
                     nbm, nbn);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

}  // namespace

extern "C" int mg_propagate(int kind, const double* meas, int64_t n_meas, const double* scat, int64_t n_scat,
                            double k0, double alpha_re, double alpha_im, int flags, double radius, const void* Ek,
                            int64_t ld_ek, int64_t nrhs, double beta_re, double beta_im, void* E, int64_t ld_e,
                            void* stream) {
  if (n_meas < 0 || n_scat < 0 || nrhs < 0) return MG_ERR_ARG;
  if (n_meas == 0 || nrhs == 0) return MG_OK;
  if (ld_ek < nrhs || ld_e < nrhs) return MG_ERR_ARG;
  if (flags & (MG_F_ZERO_SELF | MG_F_DIAG | MG_F_NANFILL)) return MG_ERR_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const double2 al = make_double2(alpha_re, alpha_im), be = make_double2(beta_re, beta_im);
  switch (kind) {
    case MG_2D_TM: return launch<MG_2D_TM>(meas, n_meas, scat, n_scat, k0, al, flags, radius, Ek, ld_ek, nrhs, be, E, ld_e, s);
    case MG_2D_TE: return launch<MG_2D_TE>(meas, n_meas, scat, n_scat, k0, al, flags, radius, Ek, ld_ek, nrhs, be, E, ld_e, s);
    case MG_3D_VECTOR: return launch<MG_3D_VECTOR>(meas, n_meas, scat, n_scat, k0, al, flags, radius, Ek, ld_ek, nrhs, be, E, ld_e, s);
    case MG_3D_SCALAR: return launch<MG_3D_SCALAR>(meas, n_meas, scat, n_scat, k0, al, flags, radius, Ek, ld_ek, nrhs, be, E, ld_e, s);
    default: return MG_ERR_KIND;
  }
}
