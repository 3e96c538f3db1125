// Complex-float GEMM C = alpha A B + beta C with every real product formed
// from bf16 MFMA products of split operands (BF16x3 or BF16x6; the tile code
// is in split_gemm.h).  They serve the mixed-precision factorisation, whose
// FP64 refinement restores full accuracy.
#include "cgemm.h"
#include "mg_internal.h"
#include "split_gemm.h"

namespace {

// TRAIL only renames the kernel (<.., 1>: the factorisation's trailing
// updates) so that profiles separate it from the other products.
template <int NS, int BM, int BN, int WM, int WN, int TRAIL>
__global__ __launch_bounds__(64 * WM * WN) void k_cgemm_split(int64_t M, int64_t N, int64_t K, float2 alpha,
                                                              const float2* __restrict__ A, int64_t lda,
                                                              const float2* __restrict__ B, int64_t ldb,
                                                              float2 beta, float2* __restrict__ C, int64_t ldc,
                                                              int nbm, int nbn) {
  typedef mg_sg::Tile<NS, BM, BN, WM, WN> Tile;
  __shared__ typename Tile::S sm;
  int tm, tn;
  mg_tile_order(blockIdx.x, nbm, nbn, tm, tn);
  const int64_t i0 = (int64_t)tm * BM, j0 = (int64_t)tn * BN;
  // The accumulators start from (beta / alpha) C, loaded before the k-loop so
  // that its latency hides behind the first k-steps; the epilogue then only
  // scales by alpha and writes.  (alpha = 0 keeps C = beta C.)
  const bool alpha0 = alpha.x == 0.f && alpha.y == 0.f;
  const bool cload = !(beta.x == 0.f && beta.y == 0.f) && !alpha0;
  float2 q = make_float2(0.f, 0.f);   // beta / alpha
  if (cload) {
    const float d = alpha.x * alpha.x + alpha.y * alpha.y;
    q = make_float2((beta.x * alpha.x + beta.y * alpha.y) / d, (beta.y * alpha.x - beta.x * alpha.y) / d);
  }
  Tile t;
  t.init(C + i0 * ldc + j0, ldc, M - i0, N - j0, q, cload);
  t.mac(sm, A + i0 * lda, lda, B + j0, ldb, M - i0, N - j0, K);
  t.store(C + i0 * ldc + j0, ldc, M - i0, N - j0, alpha, beta, alpha0);
}

template <int NS, int BM, int BN, int WM, int WN, int TRAIL>
int launch_split(int64_t M, int64_t N, int64_t K, mg::cd alpha, const void* A, int64_t lda, const void* B,
                 int64_t ldb, mg::cd beta, void* C, int64_t ldc, hipStream_t s) {
  const int nbm = (int)((M + BM - 1) / BM), nbn = (int)((N + BN - 1) / BN);
  hipLaunchKernelGGL((k_cgemm_split<NS, BM, BN, WM, WN, TRAIL>), dim3(nbm * nbn), dim3(64 * WM * WN), 0, s, M, N,
                     K,
                     make_float2((float)alpha.re, (float)alpha.im), reinterpret_cast<const float2*>(A), lda,
                     reinterpret_cast<const float2*>(B), ldb, make_float2((float)beta.re, (float)beta.im),
                     reinterpret_cast<float2*>(C), ldc, nbm, nbn);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

// 128 x 128 tiles (8 waves of 32 x 64) when the grid still gives every CU a
// tile, else 64 x 128 (4 waves): the larger tile halves the operand reads per
// flop, which bound the small one.  With three-part splits a 64 x 128 tile
// needs 90 KiB of LDS, one workgroup per CU; 64 x 64 (60 KiB) keeps two.
template <int NS, int TRAIL>
int dispatch_split(int64_t M, int64_t N, int64_t K, mg::cd alpha, const void* A, int64_t lda, const void* B,
                   int64_t ldb, mg::cd beta, void* C, int64_t ldc, hipStream_t s) {
  if (M < 0 || N < 0 || K < 0) return MG_ERR_ARG;
  if (M == 0 || N == 0) return MG_OK;
  const int64_t big = ((M + 127) / 128) * ((N + 127) / 128);
  if (big >= (int64_t)mg::cu_count())
    return launch_split<NS, 128, 128, 4, 2, TRAIL>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  if constexpr (NS == 3)
    return launch_split<NS, 64, 64, 2, 2, TRAIL>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  else
    return launch_split<NS, 64, 128, 2, 2, TRAIL>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

}  // namespace

namespace mg {

int cgemm_bf16x3(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                 cd beta, void* C, int64_t ldc, hipStream_t s) {
  return dispatch_split<2, 0>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

int cgemm_bf16x3_trail(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B,
                       int64_t ldb, cd beta, void* C, int64_t ldc, hipStream_t s) {
  return dispatch_split<2, 1>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

int cgemm_bf16x6(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                 cd beta, void* C, int64_t ldc, hipStream_t s) {
  return dispatch_split<3, 0>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

int cgemm_bf16x6_trail(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B,
                       int64_t ldb, cd beta, void* C, int64_t ldc, hipStream_t s) {
  return dispatch_split<3, 1>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

}  // namespace mg
