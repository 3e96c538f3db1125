// Complex GEMM launchers (FP64 and FP32 MFMA) and their C-ABI entry.
#include "cgemm.h"
#include "zgemm_dma.h"
#include "mg_internal.h"
#include <cstdint>
#include <cstdlib>

namespace {

// blockIdx.x walks the tile grid, blockIdx.y the batch (element strides
// sA, sB, sC between consecutive problems).
template <typename T, int BM, int BN, int BK, int WM, int WN, bool G3>
__global__ __launch_bounds__(64 * WM * WN) void k_cgemm(int64_t M, int64_t N, int64_t K,
                                                        typename Cplx<T>::t alpha, MemA a, int64_t sA,
                                                        const void* B, int64_t ldb, int64_t sB,
                                                        typename Cplx<T>::t beta, void* C, int64_t ldc, int64_t sC,
                                                        int nbm, int nbn) {
  typedef typename Cplx<T>::t T2;
  int tm, tn;
  mg_tile_order(blockIdx.x, nbm, nbn, tm, tn);
  const int64_t b = blockIdx.y;
  MemA ab{reinterpret_cast<const T2*>(a.A) + b * sA, a.lda};
  cgemm_body<T, BM, BN, BK, WM, WN, MemA, G3>(M, N, K, alpha, ab, reinterpret_cast<const T2*>(B) + b * sB, ldb,
                                              beta, reinterpret_cast<T2*>(C) + b * sC, ldc, tm, tn);
}

// The same GEMM under its own symbol for the LU trailing update, so profiles
// separate it from the small GEMMs inside panels and triangular solves.
template <typename T, int BM, int BN, int BK, int WM, int WN, bool G3>
__global__ __launch_bounds__(64 * WM * WN) void k_cgemm_trail(int64_t M, int64_t N, int64_t K,
                                                              typename Cplx<T>::t alpha, MemA a, const void* B,
                                                              int64_t ldb, typename Cplx<T>::t beta, void* C,
                                                              int64_t ldc, int nbm, int nbn) {
  int tm, tn;
  mg_tile_order(blockIdx.x, nbm, nbn, tm, tn);
  cgemm_body<T, BM, BN, BK, WM, WN, MemA, G3>(M, N, K, alpha, a, B, ldb, beta, C, ldc, tm, tn);
}

// FP64 GEMM with DMA-staged operands (zgemm_dma.h); blockIdx.y walks the
// batch as in k_cgemm.  The _trail symbol serves the LU trailing update.
template <int BN, bool LATE>
__global__ __launch_bounds__(mg_zdma::NT) void k_zgemm_dma(int64_t M, int64_t N, int64_t K, double2 alpha,
                                                          const double2* A, int64_t lda, int64_t sA,
                                                          const double2* B, int64_t ldb, int64_t sB, double2 beta,
                                                          double2* C, int64_t ldc, int64_t sC, int nbm, int nbn) {
  int tm, tn;
  mg_tile_order(blockIdx.x, nbm, nbn, tm, tn);
  const int64_t b = blockIdx.y;
  mg_zdma::zgemm_tile<BN, LATE>(M, N, K, alpha, A + b * sA, lda, B + b * sB, ldb, beta, C + b * sC, ldc, tm, tn);
}

template <int BN, bool LATE>
__global__ __launch_bounds__(mg_zdma::NT) void k_zgemm_dma_trail(int64_t M, int64_t N, int64_t K, double2 alpha,
                                                                const double2* A, int64_t lda, const double2* B,
                                                                int64_t ldb, double2 beta, double2* C, int64_t ldc,
                                                                int nbm, int nbn) {
  int tm, tn;
  mg_tile_order(blockIdx.x, nbm, nbn, tm, tn);
  mg_zdma::zgemm_tile<BN, LATE>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn);
}

template <typename T>
typename Cplx<T>::t to_t2(mg::cd z) {
  typename Cplx<T>::t r;
  r.x = (T)z.re;
  r.y = (T)z.im;
  return r;
}

using mg::cd;

// Tile configurations, measured on gfx950 for the LU trailing-update shape
// (r x r x 256) and the tall-skinny residual shape (n x T x n):
//  * 128 x 64 tiles (8 waves of 32 x 32) with three-product (Gauss) complex
//    multiplies when the grid fills the CUs: 25% fewer MFMAs, and half the
//    operand re-reads of 64 x 64 along the rows;
//  * otherwise 64 x 64 (4 waves of 32 x 32): FP64 with four products (the
//    third accumulator costs more occupancy than the saved MFMAs return at
//    this size), FP32 with three.
template <typename T, bool TRAIL, int BM, int BN, int BK, int WM, int WN, bool G3>
int launch(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, int64_t sA, const void* B,
           int64_t ldb, int64_t sB, cd beta, void* C, int64_t ldc, int64_t sC, int64_t batch, hipStream_t s) {
  const int nbm = (int)((M + BM - 1) / BM), nbn = (int)((N + BN - 1) / BN);
  MemA a{A, lda};
  if (TRAIL)
    hipLaunchKernelGGL((k_cgemm_trail<T, BM, BN, BK, WM, WN, G3>), dim3(nbm * nbn), dim3(64 * WM * WN), 0, s, M, N,
                       K, to_t2<T>(alpha), a, B, ldb, to_t2<T>(beta), C, ldc, nbm, nbn);
  else
    hipLaunchKernelGGL((k_cgemm<T, BM, BN, BK, WM, WN, G3>), dim3(nbm * nbn, (unsigned)batch), dim3(64 * WM * WN),
                       0, s, M, N, K, to_t2<T>(alpha), a, sA, B, ldb, sB, to_t2<T>(beta), C, ldc, sC, nbm, nbn);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

// MG_ZGEMM_DMA=0 selects the register-staged FP64 kernels instead (A/B timing)
inline bool zdma_enabled() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("MG_ZGEMM_DMA");
    v = (e && e[0] == '0') ? 0 : 1;
  }
  return v == 1;
}

// MG_ZGEMM_LATE=0: C loaded before the first k-tile (A/B timing)
inline bool zdma_late() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("MG_ZGEMM_LATE");
    v = (e && e[0] == '0') ? 0 : 1;
  }
  return v == 1;
}

// FP64 with DMA-staged operands: 128 x 64 tiles, three stages (one workgroup per CU)
template <bool TRAIL>
int launch_zdma(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, int64_t sA, const void* B,
                int64_t ldb, int64_t sB, cd beta, void* C, int64_t ldc, int64_t sC, int64_t batch, hipStream_t s) {
  constexpr int BN = 64;
  const int nbm = (int)((M + mg_zdma::BM - 1) / mg_zdma::BM), nbn = (int)((N + BN - 1) / BN);
  const double2 al = to_t2<double>(alpha), be = to_t2<double>(beta);
  const double2* a = reinterpret_cast<const double2*>(A);
  const double2* b = reinterpret_cast<const double2*>(B);
  double2* c = reinterpret_cast<double2*>(C);
  // C read late (zgemm_tile's LATE) when it is read at all and its elements are 16-byte aligned
  const bool late = zdma_late() && !(beta.re == 0 && beta.im == 0) && K >= 3 * mg_zdma::BK &&
                    ((reinterpret_cast<uintptr_t>(C) | (uintptr_t)(sC * 16)) & 15) == 0;
  if (TRAIL && late)
    hipLaunchKernelGGL((k_zgemm_dma_trail<BN, true>), dim3(nbm * nbn), dim3(mg_zdma::NT), 0, s, M, N, K, al, a, lda, b,
                       ldb, be, c, ldc, nbm, nbn);
  else if (TRAIL)
    hipLaunchKernelGGL((k_zgemm_dma_trail<BN, false>), dim3(nbm * nbn), dim3(mg_zdma::NT), 0, s, M, N, K, al, a, lda,
                       b, ldb, be, c, ldc, nbm, nbn);
  else if (late)
    hipLaunchKernelGGL((k_zgemm_dma<BN, true>), dim3(nbm * nbn, (unsigned)batch), dim3(mg_zdma::NT), 0, s, M, N, K,
                       al, a, lda, sA, b, ldb, sB, be, c, ldc, sC, nbm, nbn);
  else
    hipLaunchKernelGGL((k_zgemm_dma<BN, false>), dim3(nbm * nbn, (unsigned)batch), dim3(mg_zdma::NT), 0, s, M, N, K,
                       al, a, lda, sA, b, ldb, sB, be, c, ldc, sC, nbm, nbn);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

// The DMA-staged FP64 kernel takes products whose 128 x 64 tiles give every
// CU one (it holds one workgroup per CU) and whose depth is a whole number
// of k-tiles.
bool zdma_shape(int64_t M, int64_t N, int64_t K, int64_t batch) {
  const int64_t tiles = ((M + 127) / 128) * ((N + 63) / 64) * batch;
  return zdma_enabled() && tiles >= (int64_t)mg::cu_count() && K > 0 && K % mg_zdma::BK == 0 && M < (1ll << 31) &&
         N < (1ll << 31);
}

template <typename T, bool TRAIL>
int dispatch(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, int64_t sA, const void* B,
             int64_t ldb, int64_t sB, cd beta, void* C, int64_t ldc, int64_t sC, int64_t batch, hipStream_t s) {
  constexpr bool F64 = std::is_same<T, double>::value;
  const int64_t tiles = ((M + 127) / 128) * ((N + 63) / 64) * batch;
  // the DMA loads move 16-byte elements: operands at an address that is not a multiple of 16 (a
  // caller's pointer into the middle of an element) take the register-staged kernel
  const bool aligned = ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) & 15) == 0;
  if (F64 && aligned && zdma_shape(M, N, K, batch) && !(alpha.re == 0 && alpha.im == 0))
    return launch_zdma<TRAIL>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, s);
  if (tiles >= (F64 ? 1 : 2) * (int64_t)mg::cu_count())
    return launch<T, TRAIL, 128, 64, 16, 4, 2, true>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch,
                                                     s);
  return launch<T, TRAIL, 64, 64, 16, 2, 2, !F64>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch,
                                                  s);
}

// C = alpha sum_s W_s + beta C over the S split-K partials W_s (M x N, dense),
// summed in the fixed order s = 0 .. S-1 so that results do not depend on
// scheduling.
template <typename T>
__global__ __launch_bounds__(256) void k_splitk_reduce(int64_t M, int64_t N, int S, const typename Cplx<T>::t* W,
                                                       typename Cplx<T>::t alpha, typename Cplx<T>::t beta,
                                                       typename Cplx<T>::t* C, int64_t ldc) {
  typedef typename Cplx<T>::t T2;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= M * N) return;
  const int64_t i = idx / N, j = idx % N;
  T2 acc = W[idx];
  for (int q = 1; q < S; ++q) {
    const T2 w = W[(int64_t)q * M * N + idx];
    acc.x += w.x;
    acc.y += w.y;
  }
  T2 out;
  out.x = alpha.x * acc.x - alpha.y * acc.y;
  out.y = alpha.x * acc.y + alpha.y * acc.x;
  if (beta.x != 0 || beta.y != 0) {
    const T2 c = C[i * ldc + j];
    out.x += beta.x * c.x - beta.y * c.y;
    out.y += beta.x * c.y + beta.y * c.x;
  }
  C[i * ldc + j] = out;
}

// Number of K-splits for a tall-skinny product whose 128 x 64 tile grid
// leaves CUs idle: at least enough splits for two workgroups per CU, and
// preferably a number that makes the workgroup count a multiple of the CU
// count (the 128 x 64 FP64 tile holds one workgroup per CU, so a partial
// last round leaves CUs idle: 12288 x 128 x 12288 took 3.1 ms with three
// splits, 2.25 rounds); each split at least 512 deep.  1 = no split.
int splitk_parts(int64_t M, int64_t N, int64_t K) {
  const int64_t tiles = ((M + 127) / 128) * ((N + 63) / 64);
  const int64_t cus = (int64_t)mg::cu_count(), want = 2 * cus;
  if (tiles >= want || K < 1024) return 1;
  const int64_t top = std::min<int64_t>(K / 512, 16);
  int64_t S = (want + tiles - 1) / tiles;
  for (int64_t t = S; t <= top; ++t)
    if ((tiles * t) % cus == 0) return (int)t;
  S = std::min<int64_t>(S, top);
  return (int)std::max<int64_t>(S, 1);
}

}  // namespace

namespace mg {

bool zgemm_uses_dma(int64_t M, int64_t N, int64_t K, int64_t batch) { return zdma_shape(M, N, K, batch); }

int cu_count() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cache[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    cache[dev] = v;
  }
  return cache[dev];
}

template <typename T>
int cgemm_batched(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, int64_t sA, const void* B,
                  int64_t ldb, int64_t sB, cd beta, void* C, int64_t ldc, int64_t sC, int64_t batch,
                  hipStream_t s) {
  if (M < 0 || N < 0 || K < 0 || batch < 0 || batch > 65535) return MG_ERR_ARG;
  if (M == 0 || N == 0 || batch == 0) return MG_OK;
  return dispatch<T, false>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, s);
}

template <typename T>
int cgemm(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
          cd beta, void* C, int64_t ldc, hipStream_t s) {
  return cgemm_batched<T>(M, N, K, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1, s);
}

template <typename T>
int cgemm_trail(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                cd beta, void* C, int64_t ldc, hipStream_t s) {
  if (M < 0 || N < 0 || K < 0) return MG_ERR_ARG;
  if (M == 0 || N == 0) return MG_OK;
  return dispatch<T, true>(M, N, K, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1, s);
}

template <typename T>
int64_t cgemm_work_elems(int64_t M, int64_t N, int64_t K) {
  if (M <= 0 || N <= 0 || K <= 0) return 0;
  const int S = splitk_parts(M, N, K);
  return S > 1 ? (int64_t)S * M * N : 0;
}

// As cgemm, with K split over the batch dimension into a workspace of
// cgemm_work_elems(M, N, K) elements (no split when work is too small).
template <typename T>
int cgemm_ws(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
             cd beta, void* C, int64_t ldc, void* work, int64_t work_elems, hipStream_t s) {
  if (M < 0 || N < 0 || K < 0) return MG_ERR_ARG;
  if (M == 0 || N == 0) return MG_OK;
  const int S = splitk_parts(M, N, K);
  if (S <= 1 || work == nullptr || work_elems < (int64_t)S * M * N)
    return cgemm<T>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  typedef typename Cplx<T>::t T2;
  const int64_t Ks = ((K + S - 1) / S + 15) / 16 * 16;   // split depth, a multiple of the k-tile
  const int full = (int)std::min<int64_t>(S, K / Ks);     // splits of depth Ks
  const int64_t rest = K - full * Ks;                      // depth of the one shorter split (may be 0)
  const int parts = full + (rest > 0 ? 1 : 0);
  T2* W = reinterpret_cast<T2*>(work);
  const cd one{1, 0}, zero{0, 0};
  int st = dispatch<T, false>(M, N, Ks, one, A, lda, Ks, B, ldb, Ks * ldb, zero, W, N, M * N, full, s);
  if (st != MG_OK) return st;
  if (rest > 0) {
    st = dispatch<T, false>(M, N, rest, one, reinterpret_cast<const T2*>(A) + full * Ks, lda, 0,
                            reinterpret_cast<const T2*>(B) + full * Ks * ldb, ldb, 0, zero, W + (int64_t)full * M * N,
                            N, 0, 1, s);
    if (st != MG_OK) return st;
  }
  const int64_t total = M * N;
  hipLaunchKernelGGL((k_splitk_reduce<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, M, N, parts, W,
                     to_t2<T>(alpha), to_t2<T>(beta), reinterpret_cast<T2*>(C), ldc);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

#define MG_INST(T)                                                                                                 \
  template int cgemm_batched<T>(int64_t, int64_t, int64_t, cd, const void*, int64_t, int64_t, const void*,         \
                                int64_t, int64_t, cd, void*, int64_t, int64_t, int64_t, hipStream_t);              \
  template int cgemm<T>(int64_t, int64_t, int64_t, cd, const void*, int64_t, const void*, int64_t, cd, void*,      \
                        int64_t, hipStream_t);                                                                     \
  template int cgemm_trail<T>(int64_t, int64_t, int64_t, cd, const void*, int64_t, const void*, int64_t, cd,      \
                              void*, int64_t, hipStream_t);
MG_INST(double)
MG_INST(float)
template int64_t cgemm_work_elems<double>(int64_t, int64_t, int64_t);
template int64_t cgemm_work_elems<float>(int64_t, int64_t, int64_t);
template int cgemm_ws<double>(int64_t, int64_t, int64_t, cd, const void*, int64_t, const void*, int64_t, cd, void*,
                              int64_t, void*, int64_t, hipStream_t);
template int cgemm_ws<float>(int64_t, int64_t, int64_t, cd, const void*, int64_t, const void*, int64_t, cd, void*,
                             int64_t, void*, int64_t, hipStream_t);
#undef MG_INST

}  // namespace mg

extern "C" int mg_cgemm(int prec, int64_t m, int64_t n, int64_t k, double alpha_re, double alpha_im, const void* A,
                        int64_t lda, const void* B, int64_t ldb, double beta_re, double beta_im, void* C,
                        int64_t ldc, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  mg::cd al{alpha_re, alpha_im}, be{beta_re, beta_im};
  switch (prec) {
    case MG_C128: return mg::cgemm<double>(m, n, k, al, A, lda, B, ldb, be, C, ldc, s);
    case MG_C64: return mg::cgemm<float>(m, n, k, al, A, lda, B, ldb, be, C, ldc, s);
    case MG_BF16X3: return mg::cgemm_bf16x3(m, n, k, al, A, lda, B, ldb, be, C, ldc, s);
    case MG_BF16X6: return mg::cgemm_bf16x6(m, n, k, al, A, lda, B, ldb, be, C, ldc, s);
    default: return MG_ERR_ARG;
  }
}

extern "C" int64_t mg_cgemm_work_elems(int prec, int64_t m, int64_t n, int64_t k) {
  switch (prec) {
    case MG_C128: return mg::cgemm_work_elems<double>(m, n, k);
    case MG_C64: return mg::cgemm_work_elems<float>(m, n, k);
    default: return 0;
  }
}

extern "C" int mg_cgemm_ws(int prec, int64_t m, int64_t n, int64_t k, double alpha_re, double alpha_im,
                           const void* A, int64_t lda, const void* B, int64_t ldb, double beta_re, double beta_im,
                           void* C, int64_t ldc, void* work, int64_t work_elems, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  mg::cd al{alpha_re, alpha_im}, be{beta_re, beta_im};
  switch (prec) {
    case MG_C128: return mg::cgemm_ws<double>(m, n, k, al, A, lda, B, ldb, be, C, ldc, work, work_elems, s);
    case MG_C64: return mg::cgemm_ws<float>(m, n, k, al, A, lda, B, ldb, be, C, ldc, work, work_elems, s);
    default: return mg_cgemm(prec, m, n, k, alpha_re, alpha_im, A, lda, B, ldb, beta_re, beta_im, C, ldc, stream);
  }
}
