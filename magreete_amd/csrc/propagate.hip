// Fused propagation E = alpha * G(meas, scat) @ Ek + beta * E.
//
// Restates the matmul of propagate / propagate_ss (Transmission3D.py:163,
// 220; Transmission2D.py:190, 701): the (d M) x (d N) Green matrix is
// evaluated tile by tile in FP64 into LDS (GreenA loader) and contracted on
// FP64 MFMA against Ek; it never reaches HBM.  Coincident pairs evaluate to
// NaN, so the affected output rows are NaN exactly where the reference's are,
// and the caller applies the reference's row fix-up.
#include <algorithm>
#include <mutex>

#include "cgemm.h"
#include "mg_internal.h"
#include "sources.h"

namespace {

template <int KIND>
struct PropCfg {
  static constexpr int BM = 64, BN = 64, BK = 8, WM = 2, WN = 2;
};
template <>
struct PropCfg<MG_3D_VECTOR> {
  static constexpr int BM = 96, BN = 64, BK = 12, WM = 2, WN = 2;
};

// Two workgroups per CU (two waves per SIMD): one's Green evaluations run
// beside the other's MFMA work.  With parts > 1 the scatterers (K) are split
// into parts of kchunk columns: workgroup (part, tile) writes alpha G_part
// Ek_part to W + part M N (row stride N) and k_sum_parts adds them and beta E.
template <int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_propagate(
    int64_t M, int64_t N, int64_t K, double2 alpha, GreenA<KIND> g, const void* Ek, int64_t ldek, double2 beta,
    void* E, int64_t lde, int nbm, int nbn, int64_t kchunk, void* W) {
  typedef PropCfg<KIND> P;
  constexpr int D = (KIND == MG_2D_TE) ? 2 : (KIND == MG_3D_VECTOR ? 3 : 1);
  constexpr int SP = (KIND == MG_2D_TM || KIND == MG_2D_TE) ? 2 : 3;
  const int tiles = nbm * nbn;
  const int part = blockIdx.x / tiles;
  int tm, tn;
  mg_tile_order(blockIdx.x - part * tiles, nbm, nbn, tm, tn);
  if (W == nullptr) {
    cgemm_body<double, P::BM, P::BN, P::BK, P::WM, P::WN, GreenA<KIND>>(M, N, K, alpha, g, Ek, ldek, beta, E, lde,
                                                                        tm, tn);
    return;
  }
  const int64_t k0 = part * kchunk, kp = K - k0 < kchunk ? K - k0 : kchunk;
  GreenA<KIND> gp = g;
  gp.cols += (k0 / D) * SP;
  gp.ncols_pts -= k0 / D;
  cgemm_body<double, P::BM, P::BN, P::BK, P::WM, P::WN, GreenA<KIND>>(
      M, N, kp, alpha, gp, reinterpret_cast<const double2*>(Ek) + k0 * ldek, ldek, make_double2(0.0, 0.0),
      reinterpret_cast<double2*>(W) + part * M * N, N, tm, tn);
}

// E = sum over parts of W + beta E (parts in a fixed order)
__global__ __launch_bounds__(256) void k_sum_parts(int64_t M, int64_t N, int parts, const double2* __restrict__ W,
                                                   double2 beta, double2* __restrict__ E, int64_t lde) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= M * N) return;
  const int64_t i = idx / N, j = idx - i * N;
  double2 acc = W[idx];
  for (int q = 1; q < parts; ++q) {
    const double2 w = W[(int64_t)q * M * N + idx];
    acc.x += w.x;
    acc.y += w.y;
  }
  if (beta.x != 0.0 || beta.y != 0.0) {
    const double2 c = E[i * lde + j];
    acc.x += beta.x * c.x - beta.y * c.y;
    acc.y += beta.x * c.y + beta.y * c.x;
  }
  E[i * lde + j] = acc;
}

// The same sum with the incident field evaluated in place of beta E:
// E[i, t] = sum_q W_q[i, t] + E0[i, t], E0 = f_t(x_m) pol[t, a] for row
// i = d m + a (sources.h; pol null for scalar fields).  parts == 0: E already
// holds the propagated part, and the incident field is added to it.
__global__ __launch_bounds__(256) void k_sum_parts_src(int64_t M, int64_t N, int parts,
                                                       const double2* __restrict__ W, int space, int type,
                                                       const double* __restrict__ meas,
                                                       const double* __restrict__ dirs,
                                                       const double* __restrict__ pol, int d, double k0, double w,
                                                       double2* __restrict__ E, int64_t lde) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= M * N) return;
  const int64_t i = idx / N, j = idx - i * N;
  double2 acc = parts > 0 ? W[idx] : E[i * lde + j];
  for (int q = 1; q < parts; ++q) {
    const double2 v = W[(int64_t)q * M * N + idx];
    acc.x += v.x;
    acc.y += v.y;
  }
  const int64_t m = i / d;
  const int a = (int)(i - m * d);
  const mg_src_d2 e = mg_source_value(space, type, meas + space * m, dirs + space * j, k0, w);
  const double pa = pol ? pol[j * d + a] : 1.0;
  acc.x += e.x * pa;
  acc.y += e.y * pa;
  E[i * lde + j] = acc;
}

// K parts that give every CU two workgroups (at most 8, each at least 64
// k-tiles deep); kchunk a multiple of the k-tile.
template <int KIND>
void plan_parts(int64_t M, int64_t N, int64_t K, int& parts, int64_t& kchunk) {
  typedef PropCfg<KIND> P;
  const int64_t tiles = ((M + P::BM - 1) / P::BM) * ((N + P::BN - 1) / P::BN);
  const int64_t want = (2 * (int64_t)mg::cu_count() + tiles - 1) / tiles;
  const int64_t maxp = std::max<int64_t>(1, std::min<int64_t>(8, K / (64 * P::BK)));
  const int64_t p = std::max<int64_t>(1, std::min(want, maxp));
  kchunk = ((K + p - 1) / p + P::BK - 1) / P::BK * P::BK;
  parts = (int)((K + kchunk - 1) / kchunk);
}

template <int KIND>
int64_t work_bytes(int64_t n_meas, int64_t n_scat, int64_t nrhs) {
  constexpr int D = (KIND == MG_2D_TE) ? 2 : (KIND == MG_3D_VECTOR ? 3 : 1);
  const int64_t M = D * n_meas, K = D * n_scat, N = nrhs;
  int parts;
  int64_t kchunk;
  plan_parts<KIND>(M, N, K, parts, kchunk);
  return parts > 1 ? (int64_t)parts * M * N * 16 : 0;
}

template <int KIND>
int launch(const double* meas, int64_t n_meas, const double* scat, int64_t n_scat, double k0, double2 alpha,
           int flags, double radius, const void* Ek, int64_t ldek, int64_t nrhs, double2 beta, void* E, int64_t lde,
           void* work, int64_t work_bytes_, hipStream_t s) {
  typedef PropCfg<KIND> P;
  constexpr int D = (KIND == MG_2D_TE) ? 2 : (KIND == MG_3D_VECTOR ? 3 : 1);
  const int64_t M = D * n_meas, K = D * n_scat, N = nrhs;
  const int nbm = (int)((M + P::BM - 1) / P::BM), nbn = (int)((N + P::BN - 1) / P::BN);
  GreenA<KIND> g{meas, scat, n_meas, n_scat, k0, flags, radius};
  int parts = 1;
  int64_t kchunk = K;
  if (work != nullptr) plan_parts<KIND>(M, N, K, parts, kchunk);
  if (parts > 1 && work_bytes_ < (int64_t)parts * M * N * 16) parts = 1;
  if (parts == 1) {
    hipLaunchKernelGGL(k_propagate<KIND>, dim3(nbm * nbn), dim3(256), 0, s, M, N, K, alpha, g, Ek, ldek, beta, E,
                       lde, nbm, nbn, K, (void*)nullptr);
    MG_HIP(hipGetLastError());
    return MG_OK;
  }
  hipLaunchKernelGGL(k_propagate<KIND>, dim3(nbm * nbn * parts), dim3(256), 0, s, M, N, K, alpha, g, Ek, ldek, beta,
                     E, lde, nbm, nbn, kchunk, work);
  MG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((M * N + 255) / 256)), dim3(256), 0, s, M, N, parts,
                     reinterpret_cast<const double2*>(work), beta, reinterpret_cast<double2*>(E), lde);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

// E = alpha G(meas, scat) Ek + E0 with the incident field evaluated in the
// final pass (no E0 buffer is read): the split partials are summed with it,
// or, unsplit, it is added to the propagated field in place.
template <int KIND>
int launch_src(const double* meas, int64_t n_meas, const double* scat, int64_t n_scat, double k0, double2 alpha,
               int flags, double radius, const void* Ek, int64_t ldek, int64_t nrhs, int type, const double* dirs,
               const double* pol, double w, void* E, int64_t lde, void* work, int64_t work_bytes_, hipStream_t s) {
  typedef PropCfg<KIND> P;
  constexpr int D = (KIND == MG_2D_TE) ? 2 : (KIND == MG_3D_VECTOR ? 3 : 1);
  constexpr int SP = (KIND == MG_2D_TM || KIND == MG_2D_TE) ? 2 : 3;
  const int64_t M = D * n_meas, K = D * n_scat, N = nrhs;
  const int nbm = (int)((M + P::BM - 1) / P::BM), nbn = (int)((N + P::BN - 1) / P::BN);
  GreenA<KIND> g{meas, scat, n_meas, n_scat, k0, flags, radius};
  int parts = 1;
  int64_t kchunk = K;
  if (work != nullptr) plan_parts<KIND>(M, N, K, parts, kchunk);
  if (parts > 1 && work_bytes_ < (int64_t)parts * M * N * 16) parts = 1;
  if (parts == 1) {
    hipLaunchKernelGGL(k_propagate<KIND>, dim3(nbm * nbn), dim3(256), 0, s, M, N, K, alpha, g, Ek, ldek,
                       make_double2(0.0, 0.0), E, lde, nbm, nbn, K, (void*)nullptr);
  } else {
    hipLaunchKernelGGL(k_propagate<KIND>, dim3(nbm * nbn * parts), dim3(256), 0, s, M, N, K, alpha, g, Ek, ldek,
                       make_double2(0.0, 0.0), E, lde, nbm, nbn, kchunk, work);
  }
  MG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_parts_src, dim3((unsigned)((M * N + 255) / 256)), dim3(256), 0, s, M, N,
                     parts == 1 ? 0 : parts, reinterpret_cast<const double2*>(work), SP, type, meas, dirs,
                     D > 1 ? pol : nullptr, D, k0, w, reinterpret_cast<double2*>(E), lde);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

}  // namespace

namespace {

int propagate_any(int kind, const double* meas, int64_t n_meas, const double* scat, int64_t n_scat, double k0,
                  double alpha_re, double alpha_im, int flags, double radius, const void* Ek, int64_t ld_ek,
                  int64_t nrhs, double beta_re, double beta_im, void* E, int64_t ld_e, void* work,
                  int64_t work_bytes_, void* stream) {
  if (n_meas < 0 || n_scat < 0 || nrhs < 0) return MG_ERR_ARG;
  if (n_meas == 0 || nrhs == 0) return MG_OK;
  if (ld_ek < nrhs || ld_e < nrhs) return MG_ERR_ARG;
  if (flags & (MG_F_ZERO_SELF | MG_F_DIAG | MG_F_NANFILL)) return MG_ERR_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const double2 al = make_double2(alpha_re, alpha_im), be = make_double2(beta_re, beta_im);
  switch (kind) {
    case MG_2D_TM: return launch<MG_2D_TM>(meas, n_meas, scat, n_scat, k0, al, flags, radius, Ek, ld_ek, nrhs, be, E, ld_e, work, work_bytes_, s);
    case MG_2D_TE: return launch<MG_2D_TE>(meas, n_meas, scat, n_scat, k0, al, flags, radius, Ek, ld_ek, nrhs, be, E, ld_e, work, work_bytes_, s);
    case MG_3D_VECTOR: return launch<MG_3D_VECTOR>(meas, n_meas, scat, n_scat, k0, al, flags, radius, Ek, ld_ek, nrhs, be, E, ld_e, work, work_bytes_, s);
    case MG_3D_SCALAR: return launch<MG_3D_SCALAR>(meas, n_meas, scat, n_scat, k0, al, flags, radius, Ek, ld_ek, nrhs, be, E, ld_e, work, work_bytes_, s);
    default: return MG_ERR_KIND;
  }
}

}  // namespace

extern "C" int mg_propagate(int kind, const double* meas, int64_t n_meas, const double* scat, int64_t n_scat,
                            double k0, double alpha_re, double alpha_im, int flags, double radius, const void* Ek,
                            int64_t ld_ek, int64_t nrhs, double beta_re, double beta_im, void* E, int64_t ld_e,
                            void* stream) {
  return propagate_any(kind, meas, n_meas, scat, n_scat, k0, alpha_re, alpha_im, flags, radius, Ek, ld_ek, nrhs,
                       beta_re, beta_im, E, ld_e, nullptr, 0, stream);
}

extern "C" int64_t mg_propagate_work_bytes(int kind, int64_t n_meas, int64_t n_scat, int64_t nrhs) {
  if (n_meas <= 0 || n_scat <= 0 || nrhs <= 0) return 0;
  switch (kind) {
    case MG_2D_TM: return work_bytes<MG_2D_TM>(n_meas, n_scat, nrhs);
    case MG_2D_TE: return work_bytes<MG_2D_TE>(n_meas, n_scat, nrhs);
    case MG_3D_VECTOR: return work_bytes<MG_3D_VECTOR>(n_meas, n_scat, nrhs);
    case MG_3D_SCALAR: return work_bytes<MG_3D_SCALAR>(n_meas, n_scat, nrhs);
    default: return -1;
  }
}

extern "C" int mg_propagate_ws(int kind, const double* meas, int64_t n_meas, const double* scat, int64_t n_scat,
                               double k0, double alpha_re, double alpha_im, int flags, double radius, const void* Ek,
                               int64_t ld_ek, int64_t nrhs, double beta_re, double beta_im, void* E, int64_t ld_e,
                               void* work, int64_t work_bytes_, void* stream) {
  return propagate_any(kind, meas, n_meas, scat, n_scat, k0, alpha_re, alpha_im, flags, radius, Ek, ld_ek, nrhs,
                       beta_re, beta_im, E, ld_e, work, work_bytes_, stream);
}

// ---- propagation without allocation -----------------------------------------
// For callers that cannot allocate (code under graph capture, serving loops
// that forbid it): the K-split partial sums live in a fixed device array.
// Every device that runs the entry point holds its own copy of the array
// until the process ends; calls that overlap on one device share it and
// race.  Each partial is written whole before it is summed, so nothing in
// the array needs clearing between calls.
namespace {

constexpr int64_t kPropScratchBytes = 64ll << 20;
__device__ __align__(256) unsigned char scratch[kPropScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_dev[64] = {};   // scratch's device address, per device

// The address of this device's copy of scratch (looked up once per device,
// with that device current).
int scratch_address(void** out) {
  int dev = 0;
  MG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return MG_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    MG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return MG_OK;
}

}  // namespace

extern "C" int64_t mg_propagate_noalloc_max_bytes(void) { return kPropScratchBytes; }

extern "C" int mg_propagate_noalloc(int kind, const double* meas, int64_t n_meas, const double* scat,
                                    int64_t n_scat, double k0, double alpha_re, double alpha_im, int flags,
                                    double radius, const void* Ek, int64_t ld_ek, int64_t nrhs, double beta_re,
                                    double beta_im, void* E, int64_t ld_e, void* stream) {
  const int64_t need = mg_propagate_work_bytes(kind, n_meas, n_scat, nrhs);
  if (need < 0) return MG_ERR_KIND;
  if (need > kPropScratchBytes) {
    mg::set_error("propagate_noalloc: the K-split partials need %lld bytes, over the %lld-byte scratch",
                  (long long)need, (long long)kPropScratchBytes);
    return MG_ERR_ARG;
  }
  void* work = nullptr;
  if (need > 0) MG_TRY(scratch_address(&work));
  return propagate_any(kind, meas, n_meas, scat, n_scat, k0, alpha_re, alpha_im, flags, radius, Ek, ld_ek, nrhs,
                       beta_re, beta_im, E, ld_e, work, need, stream);
}

// Propagation with the plane-wave or beam incident field evaluated in the
// final pass instead of read from E: E = alpha G Ek + E0 (the same sum as
// mg_sources followed by mg_propagate_ws with beta 1).
extern "C" int mg_propagate_src(int kind, const double* meas, int64_t n_meas, const double* scat, int64_t n_scat,
                                double k0, double alpha_re, double alpha_im, int flags, double radius,
                                const void* Ek, int64_t ld_ek, int64_t nrhs, int src_type, const double* dirs,
                                const double* pol, double w, void* E, int64_t ld_e, void* work,
                                int64_t work_bytes_, void* stream) {
  if (n_meas < 0 || n_scat < 0 || nrhs < 0 || (src_type != 0 && src_type != 1)) return MG_ERR_ARG;
  if (n_meas == 0 || nrhs == 0) return MG_OK;
  if (ld_ek < nrhs || ld_e < nrhs || !dirs) return MG_ERR_ARG;
  if ((kind == MG_2D_TE || kind == MG_3D_VECTOR) && !pol) return MG_ERR_ARG;
  if (flags & (MG_F_ZERO_SELF | MG_F_DIAG | MG_F_NANFILL)) return MG_ERR_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const double2 al = make_double2(alpha_re, alpha_im);
  switch (kind) {
    case MG_2D_TM: return launch_src<MG_2D_TM>(meas, n_meas, scat, n_scat, k0, al, flags, radius, Ek, ld_ek, nrhs, src_type, dirs, pol, w, E, ld_e, work, work_bytes_, s);
    case MG_2D_TE: return launch_src<MG_2D_TE>(meas, n_meas, scat, n_scat, k0, al, flags, radius, Ek, ld_ek, nrhs, src_type, dirs, pol, w, E, ld_e, work, work_bytes_, s);
    case MG_3D_VECTOR: return launch_src<MG_3D_VECTOR>(meas, n_meas, scat, n_scat, k0, al, flags, radius, Ek, ld_ek, nrhs, src_type, dirs, pol, w, E, ld_e, work, work_bytes_, s);
    case MG_3D_SCALAR: return launch_src<MG_3D_SCALAR>(meas, n_meas, scat, n_scat, k0, al, flags, radius, Ek, ld_ek, nrhs, src_type, dirs, pol, w, E, ld_e, work, work_bytes_, s);
    default: return MG_ERR_KIND;
  }
}
