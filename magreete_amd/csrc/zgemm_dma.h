// FP64 complex GEMM tile, C = alpha A B + beta C (row-major, interleaved
// complex128), with every operand k-tile copied from memory straight to LDS
// (global_load_lds_dwordx4: lane l writes the 16 bytes of one complex
// element at the instruction's LDS base + 16 l) through three stages, so no
// registers stage the operands, no ds_write pass runs and two k-tiles are in
// flight while a third is multiplied.
//
// 128 x BN tiles (BN = 64 or 128), 8 waves of 32 x (BN / 2).  Complex
// products as three real MFMA products (Gauss: t1 = Ar Br, t2 = Ai Bi,
// t3 = (Ar + Ai)(Br + Bi); Cr = t1 - t2, Ci = t3 - t1 - t2); LDS holds the
// interleaved elements, one 128-bit read gives a lane both parts and the
// sums are formed in registers.  MFMA k-slot q of k-step s takes tile depth
// 4 q + s, so the four 16-lane groups of a fragment read elements 4 apart.
// In LDS the 16 elements of an A row are XOR-swizzled by a permutation of
// the row's low bits (16 rows, one element each, per 16-lane group), and the
// columns of B row k by the same permutation of k / 4 within each aligned run
// of 16: the reads of one fragment then hit distinct banks (the plain B rows
// gave a two-way conflict on every B read: SQ_LDS_BANK_CONFLICT a third of
// SQ_LDS_IDX_ACTIVE).
//
// Operand rows past M and columns past N are read from the last valid row /
// column (their products land in C entries that are never written); K must
// be a multiple of the k-tile (16), which the launcher checks.
#pragma once
#include "cgemm.h"

namespace mg_zdma {

constexpr int BM = 128, BK = 16, NT = 512, WM = 4, WN = 2, TM = 2;
typedef __attribute__((address_space(1))) const void* gvp;
typedef __attribute__((address_space(3))) void* lvp;

template <int BN>
struct Stage {
  double2 a[BM * BK];   // row r, element u at r BK + (u ^ swz(r))
  double2 b[BK * BN];   // row k, column c at k BN + (c ^ swz(k / 4))
};

// the A-row swizzle (the permutation measured conflict-free on gfx950 for
// 16 rows x one 16-byte unit per 16-lane group, 16 units per row)
__device__ inline int swz(int r) { return ((2 * r) & 15) ^ ((r >> 3) & 1); }

// DMA instructions per wave and k-tile: A 4 (four rows each), B BN / 32
template <int BN>
constexpr int dma_per_tile() { return 4 + BN / 32; }

// Issues k-tile kt of both operands into st.  A: the tile's first row;
// B: column j0 of row 0.  rmax / cmax: the last valid row of A / column of
// B relative to the tile.
// One 16-byte-per-lane copy global -> LDS (lane l to dst + 16 l), written
// as asm so that the compiler does not track it: the caller's s_waitcnt
// before each barrier is then the only wait (the compiler otherwise waits for
// every copy in flight before LDS reads whose stage it cannot tell apart).
__device__ __attribute__((always_inline)) inline void dma16_asm(const void* src, lvp dst) {
  const unsigned lds = __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)dst);
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(src), "s"(lds)
               : "memory", "m0");
}

template <int BN, bool ASM = false>
__device__ __attribute__((always_inline)) inline void load_tile(Stage<BN>& st, gcptr<double2> A, int64_t lda,
                                                                gcptr<double2> B, int64_t ldb, int64_t kt, int rmax,
                                                                int cmax) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t k0 = kt * BK;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r0 = wave * 16 + i * 4;
    const int r = r0 + (lane >> 4), u = (lane & 15) ^ swz(r);
    const int rr = r < rmax ? r : rmax;
    if constexpr (ASM) dma16_asm(A + (int64_t)rr * lda + k0 + u, (lvp)&st.a[r0 * BK]);
    else __builtin_amdgcn_global_load_lds((gvp)(A + (int64_t)rr * lda + k0 + u), (lvp)&st.a[r0 * BK], 16, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < BN / 32; ++i) {
    const int q = wave * (BN / 32) + i;   // 64-element run q: row q / (BN / 64), columns from 64 (q % (BN / 64))
    const int k = q / (BN / 64), c0 = (q % (BN / 64)) * 64;
    const int c = c0 + (lane ^ swz(k >> 2)), cc = c < cmax ? c : cmax;   // slot c0 + lane holds column c
    if constexpr (ASM) dma16_asm(B + (k0 + k) * ldb + cc, (lvp)&st.b[k * BN + c0]);
    else __builtin_amdgcn_global_load_lds((gvp)(B + (k0 + k) * ldb + cc), (lvp)&st.b[k * BN + c0], 16, 0, 0);
  }
}

// acc += the k-tile in st, term-major (each product term over the wave's
// TM x TN accumulators before the next), so consecutive MFMAs never update
// the same accumulator.
template <int BN, int TN>
__device__ __attribute__((always_inline)) inline void compute(const Stage<BN>& st, v4d (&t1)[TM][TN],
                                                              v4d (&t2)[TM][TN], v4d (&t3)[TM][TN]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave / WN, wn = wave % WN, li = lane & 15, lk = lane >> 4;
  const int sl = swz(li);   // rows read here are 16 m + li
  const int sb = swz(lk);   // B rows read here are 4 lk + s
  // the fragments of all four k-steps are read first, so that the MFMAs of
  // one step run while the reads of the next are in flight
  double2 a[BK / 4][TM], b[BK / 4][TN];
#pragma unroll
  for (int s = 0; s < BK / 4; ++s) {
#pragma unroll
    for (int i = 0; i < TM; ++i) a[s][i] = st.a[(wm * 32 + i * 16 + li) * BK + ((4 * lk + s) ^ sl)];
#pragma unroll
    for (int j = 0; j < TN; ++j) b[s][j] = st.b[(4 * lk + s) * BN + wn * (BN / 2) + j * 16 + (li ^ sb)];
  }
#pragma unroll
  for (int s = 0; s < BK / 4; ++s) {
    double as[TM], bs[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) as[i] = a[s][i].x + a[s][i].y;
#pragma unroll
    for (int j = 0; j < TN; ++j) bs[j] = b[s][j].x + b[s][j].y;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) t1[i][j] = mg_mfma(a[s][i].x, b[s][j].x, t1[i][j]);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) t2[i][j] = mg_mfma(a[s][i].y, b[s][j].y, t2[i][j]);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) t3[i][j] = mg_mfma(as[i], bs[j], t3[i][j]);
  }
}

// s_waitcnt immediate: at most v vector-memory operations of this wave
// outstanding (the other counters at their maximum)
constexpr unsigned vm_wait(int v) { return (unsigned)((v & 0xF) | ((v >> 4) << 14) | (0x7 << 4) | (0xF << 8)); }
constexpr unsigned kLdsWait = 0xC07F;   // lgkmcnt(0)

// One k-step of the three-stage pipeline: k-tile t is in cur; nxt receives
// k-tile t + 2 (the stage that k-tile t - 1 occupied).  Every wave has
// finished its reads of nxt before the barrier (lgkmcnt(0) after its
// compute), and k-tile t has landed for every wave after it.
// LATE: the stages freed by the last two k-tiles receive C's first two row
// parts (cpart(part, stage)); k-tile T - 1 is then waited for with the first
// part in flight (as many DMA instructions per wave as a k-tile).
template <int BN, int TN, bool LATE = false, typename CPart = int>
__device__ __attribute__((always_inline)) inline void kstep(Stage<BN>& cur, Stage<BN>& nxt, int64_t t, int64_t T,
                                                            gcptr<double2> A, int64_t lda, gcptr<double2> B,
                                                            int64_t ldb, int rmax, int cmax, v4d (&t1)[TM][TN],
                                                            v4d (&t2)[TM][TN], v4d (&t3)[TM][TN],
                                                            CPart cpart = 0) {
  if (t + 1 < T || LATE) __builtin_amdgcn_s_waitcnt(vm_wait(dma_per_tile<BN>()));   // k-tile t + 1 may be in flight
  else __builtin_amdgcn_s_waitcnt(vm_wait(0));
  __builtin_amdgcn_s_barrier();
  if constexpr (LATE) asm volatile("" ::: "memory");   // no LDS read moves above the barrier
  if (t + 2 < T) load_tile<BN, LATE>(nxt, A, lda, B, ldb, t + 2, rmax, cmax);
  else if constexpr (LATE) cpart((int)(t + 2 - T), nxt);
  compute<BN, TN>(cur, t1, t2, t3);
  __builtin_amdgcn_s_waitcnt(kLdsWait);
}

// The tile (tm, tn) of C = alpha A B + beta C; K % BK == 0, alpha != 0.
// LATE (BN = 64, K >= 3 BK, beta != 0, C 16-byte aligned): the accumulators
// start at zero and C is copied into the stages by DMA as the last k-tiles
// free them (rows 0-47 and 48-95 during the last two k-tiles' products, rows
// 96-127 after them: one complex128 element per lane, a 64-element row per
// instruction), then C = beta C + alpha acc from LDS.  With C loaded first,
// its read held up the first products.
template <int BN, bool LATE = false>
__device__ __attribute__((always_inline)) inline void zgemm_tile(int64_t M, int64_t N, int64_t K, double2 alpha,
                                                                 const void* Ap, int64_t lda, const void* Bp,
                                                                 int64_t ldb, double2 beta, void* Cp, int64_t ldc,
                                                                 int tm, int tn) {
  static_assert(!LATE || BN == 64, "C rows of 64 elements, one per lane");
  constexpr int TN = BN / 32;
  __shared__ Stage<BN> s0, s1, s2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave / WN, wn = wave % WN, li = lane & 15, lk = lane >> 4;
  const int64_t i0 = (int64_t)tm * BM, j0 = (int64_t)tn * BN;
  const int rmax = (int)(M - 1 - i0), cmax = (int)(N - 1 - j0);
  gcptr<double2> A = (gcptr<double2>)Ap + i0 * lda;
  gcptr<double2> B = (gcptr<double2>)Bp + j0;
  gptr<double2> C = (gptr<double2>)Cp;
  const int64_t T = K / BK;

  // accumulators from (beta / alpha) C: t1 = cr, t2 = 0, t3 = cr + ci, so
  // that t1 - t2 and t3 - t1 - t2 carry it; loaded before the first DMA
  const bool cload = !LATE && !(beta.x == 0 && beta.y == 0);
  double2 q;
  {
    const double d = alpha.x * alpha.x + alpha.y * alpha.y;
    q.x = (beta.x * alpha.x + beta.y * alpha.y) / d;
    q.y = (beta.y * alpha.x - beta.x * alpha.y) / d;
  }
  v4d t1[TM][TN], t2[TM][TN], t3[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int64_t col = j0 + wn * (BN / 2) + j * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = i0 + wm * 32 + i * 16 + lk + 4 * r;
        double2 c;
        c.x = 0;
        c.y = 0;
        if (cload && row < M && col < N) c = C[row * ldc + col];
        const double cr = q.x * c.x - q.y * c.y, ci = q.x * c.y + q.y * c.x;
        t1[i][j][r] = cr;
        t2[i][j][r] = 0;
        t3[i][j][r] = cr + ci;
      }
    }

  // LATE: C row part p (rows 48 p .. 48 p + 47, the last part 32 rows) into
  // stage st: wave w takes 6 (4) rows, lane l column l
  // (stages by LDS byte address: 32-bit, wave-uniform)
  auto lds_of = [](Stage<BN>& st) { return (unsigned)(uintptr_t)(lvp)&st; };
  auto cpart_at = [&](int part, unsigned base) {
    const int rpw = part < 2 ? 6 : 4;
    const int cc = lane < cmax ? lane : cmax;
#pragma unroll
    for (int e = 0; e < 6; ++e) {
      if (e < rpw) {
        const int rl = wave * rpw + e, r = 48 * part + rl, rr = r < rmax ? r : rmax;
        const unsigned lds = __builtin_amdgcn_readfirstlane(base + (unsigned)(rl * BN * 16));
        asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(
                         C + (i0 + rr) * ldc + j0 + cc),
                     "s"(lds)
                     : "memory", "m0");
      }
    }
  };
  auto cpart = [&](int part, Stage<BN>& st) { cpart_at(part, lds_of(st)); };
  if (T > 0) load_tile<BN, LATE>(s0, A, lda, B, ldb, 0, rmax, cmax);
  if (T > 1) load_tile<BN, LATE>(s1, A, lda, B, ldb, 1, rmax, cmax);
  for (int64_t t = 0; t < T; t += 3) {
    kstep<BN, TN, LATE>(s0, s2, t, T, A, lda, B, ldb, rmax, cmax, t1, t2, t3, cpart);
    if (t + 1 < T) kstep<BN, TN, LATE>(s1, s0, t + 1, T, A, lda, B, ldb, rmax, cmax, t1, t2, t3, cpart);
    if (t + 2 < T) kstep<BN, TN, LATE>(s2, s1, t + 2, T, A, lda, B, ldb, rmax, cmax, t1, t2, t3, cpart);
  }
  if constexpr (LATE) {
    // every wave is done with k-tile T - 1's stage: C's last part into it
    __builtin_amdgcn_s_barrier();
    const unsigned o0 = lds_of(s0), o1 = lds_of(s1), o2 = lds_of(s2);
    auto stage_at = [&](int64_t idx) { return idx == 0 ? o0 : (idx == 1 ? o1 : o2); };
    cpart_at(2, stage_at((T - 1) % 3));
    __builtin_amdgcn_s_waitcnt(vm_wait(0));
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int64_t cl = wn * (BN / 2) + j * 16 + li, col = j0 + cl;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int rl = wm * 32 + i * 16 + lk + 4 * r;
          const int64_t row = i0 + rl;
          const int part = rl < 48 ? 0 : (rl < 96 ? 1 : 2);
#if __HIP_DEVICE_COMPILE__
          typedef __attribute__((address_space(3))) const double* ldp;
          const ldp L = (ldp)(uintptr_t)stage_at((T - 3 + part) % 3) + 2 * ((rl - 48 * part) * BN + cl);
          const double2 c = make_double2(L[0], L[1]);
#else
          const double2 c = make_double2(0.0, 0.0);
#endif
          if (row < M && col < N) {
            const double a = t1[i][j][r], b = t2[i][j][r], d = t3[i][j][r];
            const double xr = a - b, xi = d - a - b;
            double2 out;
            out.x = beta.x * c.x - beta.y * c.y + alpha.x * xr - alpha.y * xi;
            out.y = beta.x * c.y + beta.y * c.x + alpha.x * xi + alpha.y * xr;
            C[row * ldc + col] = out;
          }
        }
      }
    return;
  }

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int64_t col = j0 + wn * (BN / 2) + j * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = i0 + wm * 32 + i * 16 + lk + 4 * r;
        if (row < M && col < N) {
          const double a = t1[i][j][r], b = t2[i][j][r], c = t3[i][j][r];
          const double xr = a - b, xi = c - a - b;
          double2 out;
          out.x = alpha.x * xr - alpha.y * xi;
          out.y = alpha.x * xi + alpha.y * xr;
          C[row * ldc + col] = out;
        }
      }
    }
}

}  // namespace mg_zdma
