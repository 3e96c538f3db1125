// Complex-float GEMM C = alpha A B + beta C with every real product formed
// from bf16 MFMA products of split operands (see bf16x3.h):
//   NS = 2 (BF16x3): x = x0 + x1, products x0 y0 + x0 y1 + x1 y0, ~16 bits;
//   NS = 3 (BF16x6): x = x0 + x1 + x2, the six products with i + j <= 2,
//                    ~24 bits (FP32-level).
// They serve the mixed-precision factorisation, whose FP64 refinement
// restores full accuracy.
//
// Tiles: BM x BN complex outputs per workgroup of WM x WN waves (each wave
// 16 TM x 16 TN), k-steps of 32 complex.  The next k-tile is fetched into
// registers while the current one is multiplied; LDS holds it split into
// 2 NS bf16 planes (re parts, im parts) with k contiguous, so each lane reads
// its 8 k-values of a plane as one 16-byte load.
#include "bf16x3.h"
#include "cgemm.h"
#include "mg_internal.h"

namespace {

constexpr int BK = 32, KP = BK + 8;   // KP: padded k stride (bf16 elements)

// x -> NS bf16 parts, each the rounded remainder of the previous ones
template <int NS>
__device__ inline void split_n(float x, unsigned short* p) {
  float r = x;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    p[i] = mg_bf16_rne(r);
    r -= __uint_as_float((unsigned)p[i] << 16);
  }
}

template <int NS, int BM, int BN>
struct alignas(16) SplitSmem {
  unsigned short a[2 * NS][BM][KP];   // planes: re part 0..NS-1, im part 0..NS-1
  unsigned short b[2 * NS][BN][KP];
};

// (r, i) += A B for one 16 x 16 x 32 complex step from split planes:
// pairs (p, q) with p + q <= NS - 1 of every real product.
template <int NS>
__device__ inline void cmac_split(const bf16x8* ar, const bf16x8* ai, const bf16x8* nai, const bf16x8* br,
                                  const bf16x8* bi, v4f& r, v4f& i) {
#pragma unroll
  for (int p = 0; p < NS; ++p)
#pragma unroll
    for (int q = 0; q + p < NS; ++q) {
      r = mg_mfma_bf16(ar[p], br[q], r);
      r = mg_mfma_bf16(nai[p], bi[q], r);
      i = mg_mfma_bf16(ar[p], bi[q], i);
      i = mg_mfma_bf16(ai[p], br[q], i);
    }
}

// TRAIL only renames the kernel (<.., 1>: the factorisation's trailing
// updates) so that profiles separate it from the other products.
template <int NS, int BM, int BN, int WM, int WN, int TRAIL>
__global__ __launch_bounds__(64 * WM * WN) void k_cgemm_split(int64_t M, int64_t N, int64_t K, float2 alpha,
                                                              const float2* __restrict__ A, int64_t lda,
                                                              const float2* __restrict__ B, int64_t ldb,
                                                              float2 beta, float2* __restrict__ C, int64_t ldc,
                                                              int nbm, int nbn) {
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
  constexpr int AROWS = NT / (BK / 8);            // A rows staged per pass (8 consecutive k per thread)
  constexpr int APASS = BM / AROWS;
  constexpr int BGRP = NT / BN;                   // k-groups of 8 staged per pass of B
  constexpr int BPASS = BK / (8 * BGRP);
  static_assert(BM % AROWS == 0 && BK % (8 * BGRP) == 0 && NT % BN == 0, "staging layout");
  __shared__ SplitSmem<NS, BM, BN> sm;
  int tm, tn;
  mg_tile_order(blockIdx.x, nbm, nbn, tm, tn);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int64_t i0 = (int64_t)tm * BM, j0 = (int64_t)tn * BN;
  const int li = lane & 15, lg = lane >> 4;
  const int ar = tid / (BK / 8), ak = (tid % (BK / 8)) * 8;
  const int bc = tid % BN, bk = (tid / BN) * 8;

  float2 ra[APASS][8], rb[BPASS][8];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int p = 0; p < APASS; ++p)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int64_t row = i0 + ar + p * AROWS, k = k0 + ak + e;
        float2 v = make_float2(0.f, 0.f);
        if (row < M && k < K) v = A[row * lda + k];
        ra[p][e] = v;
      }
#pragma unroll
    for (int p = 0; p < BPASS; ++p)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int64_t k = k0 + bk + 8 * BGRP * p + e, col = j0 + bc;
        float2 v = make_float2(0.f, 0.f);
        if (k < K && col < N) v = B[k * ldb + col];
        rb[p][e] = v;
      }
  };
  auto stash = [&](SplitSmem<NS, BM, BN>& sm) {
    unsigned short s[2 * NS][8];
#pragma unroll
    for (int p = 0; p < APASS; ++p) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        unsigned short t[NS];
        split_n<NS>(ra[p][e].x, t);
#pragma unroll
        for (int q = 0; q < NS; ++q) s[q][e] = t[q];
        split_n<NS>(ra[p][e].y, t);
#pragma unroll
        for (int q = 0; q < NS; ++q) s[NS + q][e] = t[q];
      }
#pragma unroll
      for (int q = 0; q < 2 * NS; ++q) *reinterpret_cast<u32x4*>(&sm.a[q][ar + p * AROWS][ak]) = mg_pack8(s[q]);
    }
#pragma unroll
    for (int p = 0; p < BPASS; ++p) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        unsigned short t[NS];
        split_n<NS>(rb[p][e].x, t);
#pragma unroll
        for (int q = 0; q < NS; ++q) s[q][e] = t[q];
        split_n<NS>(rb[p][e].y, t);
#pragma unroll
        for (int q = 0; q < NS; ++q) s[NS + q][e] = t[q];
      }
#pragma unroll
      for (int q = 0; q < 2 * NS; ++q)
        *reinterpret_cast<u32x4*>(&sm.b[q][bc][bk + 8 * BGRP * p]) = mg_pack8(s[q]);
    }
  };

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
  v4f cr[TM][TN], ci[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) {
      const int64_t col = j0 + wn * (BN / WN) + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = i0 + wm * (BM / WM) + a * 16 + lg * 4 + r;
        float2 c = make_float2(0.f, 0.f);
        if (cload && row < M && col < N) c = C[row * ldc + col];
        cr[a][b][r] = q.x * c.x - q.y * c.y;
        ci[a][b][r] = q.x * c.y + q.y * c.x;
      }
    }

  auto compute = [&](const SplitSmem<NS, BM, BN>& sm) {
    bf16x8 afr[TM][NS], afi[TM][NS], afn[TM][NS];
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      const int row = wm * (BM / WM) + a * 16 + li;
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        afr[a][q] = mg_ld8(&sm.a[q][row][lg * 8]);
        afi[a][q] = mg_ld8(&sm.a[NS + q][row][lg * 8]);
        afn[a][q] = mg_neg8(afi[a][q]);
      }
    }
#pragma unroll
    for (int b = 0; b < TN; ++b) {
      const int col = wn * (BN / WN) + b * 16 + li;
      bf16x8 bfr[NS], bfi[NS];
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        bfr[q] = mg_ld8(&sm.b[q][col][lg * 8]);
        bfi[q] = mg_ld8(&sm.b[NS + q][col][lg * 8]);
      }
#pragma unroll
      for (int a = 0; a < TM; ++a) cmac_split<NS>(afr[a], afi[a], afn[a], bfr, bfi, cr[a][b], ci[a][b]);
    }
  };

  const int64_t KT = (K + BK - 1) / BK;
  fetch(0);
  for (int64_t kt = 0; kt < KT; ++kt) {
    __syncthreads();   // previous tile's reads are done
    stash(sm);
    __syncthreads();
    if (kt + 1 < KT) fetch((kt + 1) * BK);
    compute(sm);
  }

#pragma unroll
  for (int a = 0; a < TM; ++a) {
#pragma unroll
    for (int b = 0; b < TN; ++b) {
      const int64_t col = j0 + wn * (BN / WN) + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = i0 + wm * (BM / WM) + a * 16 + lg * 4 + r;
        if (row < M && col < N) {
          const float xr = cr[a][b][r], xi = ci[a][b][r];
          float2 o = make_float2(alpha.x * xr - alpha.y * xi, alpha.x * xi + alpha.y * xr);
          if (alpha0) {
            const float2 c = C[row * ldc + col];
            o = make_float2(beta.x * c.x - beta.y * c.y, beta.x * c.y + beta.y * c.x);
          }
          C[row * ldc + col] = o;
        }
      }
    }
  }
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
