// Workgroup-level complex-float GEMM tile whose real products are formed from
// bf16 MFMA products of split operands (see bf16x3.h):
//   NS = 2 (BF16x3): x = x0 + x1, products x0 y0 + x0 y1 + x1 y0, ~16 bits;
//   NS = 3 (BF16x6): x = x0 + x1 + x2, the six products with i + j <= 2,
//                    ~24 bits (FP32-level).
//
// One workgroup of WM x WN waves owns a BM x BN complex accumulator tile
// (each wave 16 TM x 16 TN).  mac() adds A B over k-steps of 32 complex: the
// next k-tile is fetched into registers while the current one is multiplied;
// LDS holds it split into 2 NS bf16 planes (re parts, im parts) with k
// contiguous, so each lane reads its 8 k-values of a plane as one 16-byte
// load.  fetch / stash / compute are its steps, for loops that chain k-tiles
// of several products (one k-loop keeps register pressure at one GEMM's).
#pragma once
#include "bf16x3.h"

namespace mg_sg {

// KP: padded k stride (bf16 elements): 96-byte rows.  On gfx950 the 128-bit
// reads of a fragment (16 rows, one 16-byte unit each) are free of bank
// conflicts at a 96-byte row stride; 80 bytes (BK + 8) gave a two-way conflict
// on every read (LDS probe: 232 us against 142 us, SQ_LDS_BANK_CONFLICT half
// of the LDS-active cycles), as did 48, 112 and 144.
constexpr int BK = 32, KP = BK + 16;

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

// One padding row per plane shifts consecutive planes by 24 LDS banks, so
// that lanes storing the same (row, k) of different planes do not collide.
template <int NS, int BM, int BN>
struct alignas(16) Smem {
  unsigned short a[2 * NS][BM + 1][KP];   // planes: re part 0..NS-1, im part 0..NS-1
  unsigned short b[2 * NS][BN + 1][KP];
};

// (r, i) += A B for one 16 x 16 x 32 complex step from split planes:
// pairs (p, q) with p + q <= NS - 1 of every real product.
template <int NS>
__device__ inline void cmac(const bf16x8* ar, const bf16x8* ai, const bf16x8* nai, const bf16x8* br,
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

// Operands and accumulators live in device memory: the tile reads and writes
// them through the global address space, also when a pointer was read from
// memory (a grouped launch's descriptor table), where the compiler would
// otherwise emit flat accesses, which also count against the LDS counter (the
// waits for a k-tile's LDS reads would then wait for the next k-tile's loads).
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(1))) float2* gcf2;
typedef __attribute__((address_space(1))) float2* gf2;
#else
typedef const float2* gcf2;   // host pass: never executed
typedef float2* gf2;
#endif

// Agent-scope (cross-XCD coherent) element access, for data that other
// workgroups of the same launch write.
__device__ inline float2 coherent_load(const float2* p) {
  const unsigned long long u =
      __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __builtin_bit_cast(float2, u);
}
__device__ inline void coherent_store(float2* p, float2 v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), __builtin_bit_cast(unsigned long long, v),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int NS, int BM, int BN, int WM, int WN>
struct Tile {
  static constexpr int NT = 64 * WM * WN;
  static constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
  static constexpr int AROWS = NT / (BK / 8);   // A rows staged per pass (8 consecutive k per thread)
  static constexpr int APASS = BM / AROWS;
  static constexpr int BGRP = NT / BN;          // k-groups of 8 staged per pass of B
  static constexpr int BPASS = BK / (8 * BGRP);
  typedef Smem<NS, BM, BN> S;
  // the complex-float staging of fetch / stash (pre-split operands stage their own way)
  static constexpr bool kFloatStaging = BM % AROWS == 0 && BK % (8 * BGRP) == 0 && NT % BN == 0;

  v4f cr[TM][TN], ci[TM][TN];

  // Thread index within the tile's NT threads (a workgroup may hold several
  // tiles' thread groups, each NT consecutive threads).
  __device__ static int ltid() { return (int)threadIdx.x % NT; }

  // Tile row / column of accumulator item (a, b, r) of this lane.
  __device__ static int row_of(int a, int r) {
    const int lane = ltid() & 63, wave = ltid() >> 6;
    return (wave / WN) * (BM / WM) + a * 16 + (lane >> 4) * 4 + r;
  }
  __device__ static int col_of(int b) {
    const int lane = ltid() & 63, wave = ltid() >> 6;
    return (wave % WN) * (BN / WN) + b * 16 + (lane & 15);
  }

  // acc = q C over the valid M x N corner of C (at the tile origin), 0
  // elsewhere and everywhere when !load.
  // All loads are issued before any is used (a whole tile inside C takes the
  // branch-free path), so that they overlap.
  __device__ void init(const float2* Cp, int64_t ldc, int64_t M, int64_t N, float2 q, bool load) {
    const gcf2 C = (gcf2)Cp;
    float2 c[TM][TN][4];
    const bool whole = load && M >= BM && N >= BN;
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b) {
        const int col = col_of(b);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = row_of(a, r);
          if (whole) {
            c[a][b][r] = C[row * ldc + col];
          } else {
            c[a][b][r] = make_float2(0.f, 0.f);
            if (load && row < M && col < N) c[a][b][r] = C[row * ldc + col];
          }
        }
      }
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          cr[a][b][r] = q.x * c[a][b][r].x - q.y * c[a][b][r].y;
          ci[a][b][r] = q.x * c[a][b][r].y + q.y * c[a][b][r].x;
        }
  }

  // Operand k-tile in flight (registers), as fetched from global memory.
  float2 ra[APASS][8], rb[BPASS][8];

  // Fetch k-tile [k0, k0 + 32) of A (M x K, lda) and B (K x N, ldb), both at
  // the tile origin; entries outside M, N, K read as 0.  COHERENT_B: B with
  // agent-scope loads (B written by other workgroups of this launch).
  template <bool COHERENT_B = false>
  __device__ void fetch(const float2* __restrict__ Ap, int64_t lda, const float2* B, int64_t ldb, int64_t M,
                        int64_t N, int64_t K, int64_t k0) {
    static_assert(kFloatStaging, "staging layout");
    const gcf2 A = (gcf2)Ap, Bg = (gcf2)B;
    const int tid = ltid();
    const int ar = tid / (BK / 8), ak = (tid % (BK / 8)) * 8;
    const int bc = tid % BN, bk = (tid / BN) * 8;
    const bool kin = K - k0 >= BK;   // the whole k-tile is inside K
    const bool ain = kin && M >= BM, bin = kin && N >= BN;
#pragma unroll
    for (int p = 0; p < APASS; ++p)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int64_t row = ar + p * AROWS, k = k0 + ak + e;
        if (ain) {
          ra[p][e] = A[row * lda + k];
        } else {
          ra[p][e] = make_float2(0.f, 0.f);
          if (row < M && k < K) ra[p][e] = A[row * lda + k];
        }
      }
#pragma unroll
    for (int p = 0; p < BPASS; ++p)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int64_t k = k0 + bk + 8 * BGRP * p + e, col = bc;
        if (bin || (k < K && col < N)) {
          if constexpr (COHERENT_B) rb[p][e] = coherent_load(&B[k * ldb + col]);
          else rb[p][e] = Bg[k * ldb + col];
        } else {
          rb[p][e] = make_float2(0.f, 0.f);
        }
      }
  }

  // The fetched k-tile, split into bf16 planes, into LDS.
  __device__ void stash(S& s) const {
    static_assert(kFloatStaging, "staging layout");
    const int tid = ltid();
    const int ar = tid / (BK / 8), ak = (tid % (BK / 8)) * 8;
    const int bc = tid % BN, bk = (tid / BN) * 8;
    unsigned short t2[2 * NS][8];
#pragma unroll
    for (int p = 0; p < APASS; ++p) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        unsigned short t[NS];
        split_n<NS>(ra[p][e].x, t);
#pragma unroll
        for (int q = 0; q < NS; ++q) t2[q][e] = t[q];
        split_n<NS>(ra[p][e].y, t);
#pragma unroll
        for (int q = 0; q < NS; ++q) t2[NS + q][e] = t[q];
      }
#pragma unroll
      for (int q = 0; q < 2 * NS; ++q) *reinterpret_cast<u32x4*>(&s.a[q][ar + p * AROWS][ak]) = mg_pack8(t2[q]);
    }
#pragma unroll
    for (int p = 0; p < BPASS; ++p) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        unsigned short t[NS];
        split_n<NS>(rb[p][e].x, t);
#pragma unroll
        for (int q = 0; q < NS; ++q) t2[q][e] = t[q];
        split_n<NS>(rb[p][e].y, t);
#pragma unroll
        for (int q = 0; q < NS; ++q) t2[NS + q][e] = t[q];
      }
#pragma unroll
      for (int q = 0; q < 2 * NS; ++q) *reinterpret_cast<u32x4*>(&s.b[q][bc][bk + 8 * BGRP * p]) = mg_pack8(t2[q]);
    }
  }

  // acc += the k-tile in LDS.
  __device__ void compute(const S& s) {
    const int lane = ltid() & 63, wave = ltid() >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int li = lane & 15, lg = lane >> 4;
    bf16x8 afr[TM][NS], afi[TM][NS], afn[TM][NS];
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      const int row = wm * (BM / WM) + a * 16 + li;
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        afr[a][q] = mg_ld8(&s.a[q][row][lg * 8]);
        afi[a][q] = mg_ld8(&s.a[NS + q][row][lg * 8]);
        afn[a][q] = mg_neg8(afi[a][q]);
      }
    }
    // Term-major over pairs of column blocks (BF16x3; BF16x6 holds one block's
    // six B fragments, pairs would spill): every product term goes over the
    // group's accumulators before the next term, so consecutive MFMAs update
    // different accumulators (each accumulator still takes its terms in
    // cmac's order: the same bits as block by block).
    constexpr int HB = (TN % 2 == 0 && NS == 2) ? 2 : 1;
#pragma unroll
    for (int h = 0; h < TN; h += HB) {
      bf16x8 bfr[HB][NS], bfi[HB][NS];
#pragma unroll
      for (int b = 0; b < HB; ++b) {
        const int col = wn * (BN / WN) + (h + b) * 16 + li;
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          bfr[b][q] = mg_ld8(&s.b[q][col][lg * 8]);
          bfi[b][q] = mg_ld8(&s.b[NS + q][col][lg * 8]);
        }
      }
#pragma unroll
      for (int p = 0; p < NS; ++p)
#pragma unroll
        for (int q = 0; q + p < NS; ++q) {
#pragma unroll
          for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < HB; ++b) cr[a][h + b] = mg_mfma_bf16(afr[a][p], bfr[b][q], cr[a][h + b]);
#pragma unroll
          for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < HB; ++b) cr[a][h + b] = mg_mfma_bf16(afn[a][p], bfi[b][q], cr[a][h + b]);
#pragma unroll
          for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < HB; ++b) ci[a][h + b] = mg_mfma_bf16(afr[a][p], bfi[b][q], ci[a][h + b]);
#pragma unroll
          for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < HB; ++b) ci[a][h + b] = mg_mfma_bf16(afi[a][p], bfr[b][q], ci[a][h + b]);
        }
    }
  }

  // acc += A B with A (M x K, lda) and B (K x N, ldb) given at the tile
  // origin; entries outside M, N, K count as 0.
  __device__ void mac(S& sm, const float2* __restrict__ A, int64_t lda, const float2* __restrict__ B, int64_t ldb,
                      int64_t M, int64_t N, int64_t K) {
    const int64_t KT = (K + BK - 1) / BK;
    fetch(A, lda, B, ldb, M, N, K, 0);
    for (int64_t kt = 0; kt < KT; ++kt) {
      __syncthreads();   // previous tile's reads are done
      stash(sm);
      __syncthreads();
      if (kt + 1 < KT) fetch(A, lda, B, ldb, M, N, K, (kt + 1) * BK);
      compute(sm);
    }
  }

  // C = alpha acc over the valid M x N corner (alpha0: C = beta C instead).
  // COHERENT: agent-scope stores (read by other workgroups of this launch).
  template <bool COHERENT = false>
  __device__ void store(float2* C, int64_t ldc, int64_t M, int64_t N, float2 alpha, float2 beta, bool alpha0) const {
    const gf2 Cg = (gf2)C;
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b) {
        const int col = col_of(b);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = row_of(a, r);
          if (row < M && col < N) {
            const float xr = cr[a][b][r], xi = ci[a][b][r];
            float2 o = make_float2(alpha.x * xr - alpha.y * xi, alpha.x * xi + alpha.y * xr);
            if (alpha0) {
              const float2 c = Cg[row * ldc + col];
              o = make_float2(beta.x * c.x - beta.y * c.y, beta.x * c.y + beta.y * c.x);
            }
            if constexpr (COHERENT) coherent_store(&C[row * ldc + col], o);
            else Cg[row * ldc + col] = o;
          }
        }
      }
  }
};

}  // namespace mg_sg
