// Complex-float GEMM C = alpha A B + beta C with every real product formed
// as three bf16 MFMA products (BF16x3, see bf16x3.h).  It serves the
// mixed-precision factorisation, whose FP64 refinement restores accuracy.
//
// Tiles: 64 x 128 complex outputs per 256-thread workgroup (2 x 2 waves of
// 32 x 64, i.e. 2 x 4 MFMA tiles each), k-steps of 32 complex.  The next
// k-tile is fetched into registers while the current one is multiplied; LDS
// holds it split into four bf16 planes (re hi, re lo, im hi, im lo) with k
// contiguous, so each lane reads its 8 k-values as one 16-byte load.
#include "bf16x3.h"
#include "cgemm.h"
#include "mg_internal.h"

namespace {

constexpr int BM = 64, BN = 128, BK = 32, KP = BK + 8;   // KP: padded k stride (bf16 elements)

struct alignas(16) Smem {
  unsigned short a[4][BM][KP];   // planes: re hi, re lo, im hi, im lo
  unsigned short b[4][BN][KP];
};

// TRAIL only renames the kernel (k_cgemm_bf16x3<1>: the factorisation's
// trailing updates) so that profiles separate it from the other products.
template <int TRAIL>
__global__ __launch_bounds__(256) void k_cgemm_bf16x3(int64_t M, int64_t N, int64_t K, float2 alpha,
                                                      const float2* __restrict__ A, int64_t lda,
                                                      const float2* __restrict__ B, int64_t ldb, float2 beta,
                                                      float2* __restrict__ C, int64_t ldc, int nbm, int nbn) {
  __shared__ Smem sm;
  int tm, tn;
  mg_tile_order(blockIdx.x, nbm, nbn, tm, tn);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t i0 = (int64_t)tm * BM, j0 = (int64_t)tn * BN;
  const int li = lane & 15, lg = lane >> 4;

  // A tile: thread -> row tid/4, k-group (tid%4)*8 (8 consecutive complex = 64 B)
  const int ar = tid >> 2, ak = (tid & 3) * 8;
  // B tile: thread -> column tid%128, k-groups (tid/128)*8 and +16 (8 consecutive k each)
  const int bc = tid & 127, bk = (tid >> 7) * 8;

  float2 ra[8], rb[16];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t row = i0 + ar, k = k0 + ak + e;
      float2 v = make_float2(0.f, 0.f);
      if (row < M && k < K) v = A[row * lda + k];
      ra[e] = v;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int64_t k = k0 + bk + 16 * h + e, col = j0 + bc;
        float2 v = make_float2(0.f, 0.f);
        if (k < K && col < N) v = B[k * ldb + col];
        rb[h * 8 + e] = v;
      }
    }
  };
  auto stash = [&]() {
    unsigned short p[4][8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      mg_split(ra[e].x, p[0][e], p[1][e]);
      mg_split(ra[e].y, p[2][e], p[3][e]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) *reinterpret_cast<u32x4*>(&sm.a[q][ar][ak]) = mg_pack8(p[q]);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        mg_split(rb[h * 8 + e].x, p[0][e], p[1][e]);
        mg_split(rb[h * 8 + e].y, p[2][e], p[3][e]);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) *reinterpret_cast<u32x4*>(&sm.b[q][bc][bk + 16 * h]) = mg_pack8(p[q]);
    }
  };

  v4f cr[2][4], ci[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) { cr[a][b][r] = 0.f; ci[a][b][r] = 0.f; }

  const int64_t KT = (K + BK - 1) / BK;
  fetch(0);
  for (int64_t kt = 0; kt < KT; ++kt) {
    __syncthreads();   // previous tile's reads are done
    stash();
    __syncthreads();
    if (kt + 1 < KT) fetch((kt + 1) * BK);
    bf16x8 arh[2], arl[2], aih[2], ail[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int row = wm * 32 + a * 16 + li;
      arh[a] = mg_ld8(&sm.a[0][row][lg * 8]);
      arl[a] = mg_ld8(&sm.a[1][row][lg * 8]);
      aih[a] = mg_ld8(&sm.a[2][row][lg * 8]);
      ail[a] = mg_ld8(&sm.a[3][row][lg * 8]);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int col = wn * 64 + b * 16 + li;
      const bf16x8 brh = mg_ld8(&sm.b[0][col][lg * 8]);
      const bf16x8 brl = mg_ld8(&sm.b[1][col][lg * 8]);
      const bf16x8 bih = mg_ld8(&sm.b[2][col][lg * 8]);
      const bf16x8 bil = mg_ld8(&sm.b[3][col][lg * 8]);
#pragma unroll
      for (int a = 0; a < 2; ++a) mg_cmac12(arh[a], arl[a], aih[a], ail[a], brh, brl, bih, bil, cr[a][b], ci[a][b]);
    }
  }

  const bool beta0 = beta.x == 0.f && beta.y == 0.f;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t col = j0 + wn * 64 + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = i0 + wm * 32 + a * 16 + lg * 4 + r;
        if (row < M && col < N) {
          const float xr = cr[a][b][r], xi = ci[a][b][r];
          float2 o = make_float2(alpha.x * xr - alpha.y * xi, alpha.x * xi + alpha.y * xr);
          if (!beta0) {
            const float2 c = C[row * ldc + col];
            o.x += beta.x * c.x - beta.y * c.y;
            o.y += beta.x * c.y + beta.y * c.x;
          }
          C[row * ldc + col] = o;
        }
      }
    }
  }
}

}  // namespace

namespace mg {

template <int TRAIL>
static int launch_bf16x3(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B,
                         int64_t ldb, cd beta, void* C, int64_t ldc, hipStream_t s) {
  if (M < 0 || N < 0 || K < 0) return MG_ERR_ARG;
  if (M == 0 || N == 0) return MG_OK;
  const int nbm = (int)((M + BM - 1) / BM), nbn = (int)((N + BN - 1) / BN);
  hipLaunchKernelGGL(k_cgemm_bf16x3<TRAIL>, dim3(nbm * nbn), dim3(256), 0, s, M, N, K,
                     make_float2((float)alpha.re, (float)alpha.im), reinterpret_cast<const float2*>(A), lda,
                     reinterpret_cast<const float2*>(B), ldb, make_float2((float)beta.re, (float)beta.im),
                     reinterpret_cast<float2*>(C), ldc, nbm, nbn);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

int cgemm_bf16x3(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                 cd beta, void* C, int64_t ldc, hipStream_t s) {
  return launch_bf16x3<0>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

int cgemm_bf16x3_trail(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B,
                       int64_t ldb, cd beta, void* C, int64_t ldc, hipStream_t s) {
  return launch_bf16x3<1>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

}  // namespace mg
