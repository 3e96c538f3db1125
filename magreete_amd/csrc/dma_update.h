// Split-bf16 (BF16x3) complex update of one 128 x 128 tile, C -= A B, from
// pre-split operands (psplit.h's format: four 16-byte units per row or
// column and group of 8 k), with every operand k-tile copied from memory
// straight to LDS (global_load_lds_dwordx4: lane l writes 16 bytes at the
// instruction's LDS base + 16 l), two stages, so no registers stage the
// operands and no ds_write pass runs.
//
// 8 waves of 32 x 64 (two 16-row fragments by four 16-column fragments).
// In LDS the 16 units of a row are XOR-swizzled by a permutation of the
// row's low bits (swz) so that the reads of one fragment hit distinct banks.
// The accumulators start from -C, loaded behind the first k-tile's operands,
// or (LATE) from zero with C copied to LDS by DMA during the last k-tiles;
// the k-loop is unrolled (the tile count is a template parameter) so that
// the compiler can tell the two stages' DMA apart and waits only where the
// loop asks for it.
//
// Used by the single-problem LDU trailing update (ldu.hip) and the grouped
// LU's trailing update (group.hip); the caller supplies the k-tile source.
#pragma once
#include "bf16x3.h"

namespace mg_dma {

constexpr int BM = 128, BN = 128, WN = 2, NT = 512;
constexpr int U = 4, UPR = 16, TM = 2, TN = 4;
struct Stage {
  u32x4 a[BM * UPR];
  u32x4 b[BN * UPR];
};
typedef __attribute__((address_space(1))) const void* gvp;
typedef __attribute__((address_space(3))) void* lvp;

// LDS slot of 16-byte unit u in row r is u ^ swz(r).  The permutation
// (2 r mod 16) ^ (r / 8 mod 2) of the row's low bits makes the 128-bit reads
// of compute() (16 rows, one unit each, per 16-lane group) free of bank
// conflicts on gfx950; the plain u ^ (r mod 16) gave a two-way conflict on
// every read (LDS probe: SQ_LDS_BANK_CONFLICT equal to half the LDS-active
// cycles, 237 us against 188 us for the same reads).
__device__ inline int swz(int r) { return ((2 * r) & 15) ^ ((r >> 3) & 1); }

// Rows [16 w, 16 w + 16) of both operands' k-tile (groups g0 .. g0 + 3) for
// wave w: four DMA instructions per operand, each four whole rows.  A and B
// point at the tile's first row / column; ldg: groups per row / column.
__device__ inline void load_tile(Stage& st, const u32x4* A, const u32x4* B, int64_t ldg, int64_t g0) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r0 = wave * 16 + i * 4;
    const int r = r0 + (lane >> 4), u = (lane & 15) ^ swz(r);
    __builtin_amdgcn_global_load_lds((gvp)(A + ((int64_t)r * ldg + g0) * U + u), (lvp)&st.a[r0 * UPR], 16, 0, 0);
    __builtin_amdgcn_global_load_lds((gvp)(B + ((int64_t)r * ldg + g0) * U + u), (lvp)&st.b[r0 * UPR], 16, 0, 0);
  }
}

__device__ inline bf16x8 unit(const u32x4* p) { return __builtin_bit_cast(bf16x8, *p); }

// acc += the k-tile in st (BF16x3: part pairs (p, q) with p + q <= 1 of every
// real product, the imaginary parts of B negated for the real accumulator).
// Term-major: for each half of the wave's columns, every product term goes
// over that half's eight accumulators before the next term, so consecutive
// MFMAs update different accumulators (the block-by-block order chained six
// dependent MFMAs per accumulator; standalone 12032^2 x 256: 0.97 ms against
// 1.18, the same bits: every accumulator sees its terms in the same order).
__device__ inline void compute(const Stage& st, v4f (&cr)[TM][TN], v4f (&ci)[TM][TN]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave / WN, wn = wave % WN, li = lane & 15, lg = lane >> 4;
  const int sl = swz(li);   // rows and columns read here are 16 k + li
  bf16x8 ar[TM][2], ai[TM][2];
#pragma unroll
  for (int a = 0; a < TM; ++a) {
    const int row = wm * 32 + a * 16 + li;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      ar[a][q] = unit(&st.a[row * UPR + ((lg * U + q) ^ sl)]);
      ai[a][q] = unit(&st.a[row * UPR + ((lg * U + 2 + q) ^ sl)]);
    }
  }
#pragma unroll
  for (int h = 0; h < TN; h += 2) {
    bf16x8 br[2][2], bi[2][2], bn[2][2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int col = wn * 64 + (h + b) * 16 + li;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        br[b][q] = unit(&st.b[col * UPR + ((lg * U + q) ^ sl)]);
        bi[b][q] = unit(&st.b[col * UPR + ((lg * U + 2 + q) ^ sl)]);
        bn[b][q] = mg_neg8(bi[b][q]);
      }
    }
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = 0; q + p < 2; ++q) {
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) cr[a][h + b] = mg_mfma_bf16(ar[a][p], br[b][q], cr[a][h + b]);
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) ci[a][h + b] = mg_mfma_bf16(ar[a][p], bi[b][q], ci[a][h + b]);
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) cr[a][h + b] = mg_mfma_bf16(ai[a][p], bn[b][q], cr[a][h + b]);
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) ci[a][h + b] = mg_mfma_bf16(ai[a][p], br[b][q], ci[a][h + b]);
      }
  }
}

// s_waitcnt immediate that waits until at most v vector-memory operations
// of this wave are outstanding (other counters left at their maximum).
constexpr unsigned vm_wait(int v) { return (unsigned)((v & 0xF) | ((v >> 4) << 14) | (0x7 << 4) | (0xF << 8)); }
constexpr unsigned kLdsWait = 0xC07F;   // lgkmcnt(0)

// Rows [64 h, 64 h + 64) of the C tile into stage st (row-major, 128
// complex floats per row): one DMA instruction per row, 8 per wave.  Rows
// past M read row M - 1 and units past N read column 0 (never used).  Needs
// 16-byte aligned units: C and ldc even in complex floats, N even.
__device__ inline void load_c_half(Stage& st, const float2* C, int64_t ldc, int64_t M, int64_t N, int h) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float2* L = reinterpret_cast<float2*>(&st);
  const int64_t c = 2 * lane, cc = c < N ? c : 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int r = wave * 8 + i;
    const int64_t gr = 64 * h + r, rr = gr < M ? gr : M - 1;
    __builtin_amdgcn_global_load_lds((gvp)(C + rr * ldc + cc), (lvp)&L[r * BN], 16, 0, 0);
  }
}

// C -= sum over k-tiles t < T of the tile's products; src(t, stage) issues
// k-tile t's DMA (load_tile).  C points at the tile's origin (row-major,
// ldc); M x N: the valid extent from there (the whole tile when both are at
// least 128).  s0, s1: the workgroup's two LDS stages.  C is accessed
// through the global address space: a pointer read from memory (a grouped
// launch's descriptor table) would otherwise give flat accesses, which also
// count against the LDS counter, so the k-loop's waits for its LDS reads
// would wait on them too (measured: 199 against 305 TFLOP/s per launch).
//
// LATE = false: the accumulators start from -C, loaded behind the first
// k-tile's operands, and C = -acc at the end.  LATE = true: they start from
// zero, and C is copied by DMA into each stage as soon as its last k-tile has
// been read (the first half during the last k-tile's products), then
// C = C - acc from LDS: the tile's C read no longer holds up its first
// products (standalone, 12288^2 x 512: the C read and write took a third of
// the launch).  LATE needs load_c_half's alignment.
template <int T, bool LATE = false, typename Src>
__device__ inline void tile_update(Stage& s0, Stage& s1, Src src, float2* __restrict__ Cflat, int64_t ldc, int64_t M,
                                   int64_t N) {
  static_assert(T % 2 == 0 && T > 0, "k-tiles come in pairs (one per stage)");
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(1))) float2* gfloat2p;
#else
  typedef float2* gfloat2p;   // host pass: never executed
#endif
  const gfloat2p C = (gfloat2p)Cflat;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave / WN, wn = wave % WN;
  v4f cr[TM][TN], ci[TM][TN];
  const bool whole = M >= BM && N >= BN;
  const int64_t rbase = wm * 32 + (lane >> 4) * 4, cbase = wn * 64 + (lane & 15);
  src(0, s0);
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float2 c = make_float2(0.f, 0.f);
        if (!LATE) {
          const int64_t row = rbase + a * 16 + r, col = cbase + b * 16;
          if (whole || (row < M && col < N)) c = C[row * ldc + col];
        }
        cr[a][b][r] = -c.x;
        ci[a][b][r] = -c.y;
      }
  src(1, s1);
  // bare s_barrier (no fence): the waits before it order the DMA and the reads
#pragma unroll
  for (int t = 0; t < T; t += 2) {
    __builtin_amdgcn_s_waitcnt(vm_wait(8));     // k-tile t has landed (t + 1 may be in flight)
    __builtin_amdgcn_s_barrier();
    compute(s0, cr, ci);
    __builtin_amdgcn_s_waitcnt(kLdsWait);       // this wave's reads of s0 are done
    __builtin_amdgcn_s_barrier();
    if (t + 2 < T) {
      src(t + 2, s0);
      __builtin_amdgcn_s_waitcnt(vm_wait(8));   // k-tile t + 1 has landed
    } else if (LATE) {
      load_c_half(s0, Cflat, ldc, M, N, 0);
      __builtin_amdgcn_s_waitcnt(vm_wait(8));   // k-tile t + 1 has landed (C's first half may be in flight)
    } else {
      __builtin_amdgcn_s_waitcnt(vm_wait(0));
    }
    __builtin_amdgcn_s_barrier();
    compute(s1, cr, ci);
    __builtin_amdgcn_s_waitcnt(kLdsWait);
    __builtin_amdgcn_s_barrier();
    if (t + 3 < T) src(t + 3, s1);
  }
  if (LATE) {
    load_c_half(s1, Cflat, ldc, M, N, 1);
    __builtin_amdgcn_s_waitcnt(vm_wait(0));
    __builtin_amdgcn_s_barrier();
    const float2* L = reinterpret_cast<const float2*>(wm < 2 ? &s0 : &s1);
    const int64_t lbase = rbase - (wm < 2 ? 0 : 64);
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = rbase + a * 16 + r, col = cbase + b * 16;
          const float2 c = L[(lbase + a * 16 + r) * BN + col];
          if (whole || (row < M && col < N)) C[row * ldc + col] = make_float2(c.x - cr[a][b][r], c.y - ci[a][b][r]);
        }
    return;
  }
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = rbase + a * 16 + r, col = cbase + b * 16;
        if (whole || (row < M && col < N)) C[row * ldc + col] = make_float2(-cr[a][b][r], -ci[a][b][r]);
      }
}

}  // namespace mg_dma
