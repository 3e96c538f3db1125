// Pre-split operands for the split-bf16 complex products.
//
// The split GEMM tile (split_gemm.h) converts every float operand it loads
// into NS bf16 parts.  An operand that several products read (the LDU
// factor's panels, the right-hand sides of a sweep) is cheaper to split once
// and store split: the products then copy 16-byte units into LDS and spend no
// VALU work on conversions.
//
// Format: for each row r of an A operand (k along the row), or each column c
// of a B operand (k down the column), and each group g of 8 consecutive k,
// the 2 NS planes of those 8 elements follow each other as 16-byte units:
//   unit ((r * ldg + g) * 2 NS + q),  q < NS: real part q, q >= NS: imaginary
//   part q - NS,
// each part the bf16 rounding of what the earlier parts leave.  ldg (groups
// per row or column) is a multiple of 4, so a k-tile of 32 is 4 groups; k
// past the operand's extent holds zeros.
#pragma once
#include "split_gemm.h"

namespace mg_ps {

using mg_sg::BK;

// Eight complex values -> their 2 NS units.
template <int NS>
__device__ inline void split8(const float2* v, u32x4* out) {
  unsigned short t[2 * NS][8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    unsigned short p[NS];
    mg_sg::split_n<NS>(v[e].x, p);
#pragma unroll
    for (int q = 0; q < NS; ++q) t[q][e] = p[q];
    mg_sg::split_n<NS>(v[e].y, p);
#pragma unroll
    for (int q = 0; q < NS; ++q) t[NS + q][e] = p[q];
  }
#pragma unroll
  for (int q = 0; q < 2 * NS; ++q) out[q] = mg_pack8(t[q]);
}

// A 16-byte unit read through the global address space.  Pointers that
// reach a load through a run-time-indexed array lose their address space, and
// the compiler then emits flat loads, which also count against the LDS
// counter: every wait for the tile's LDS reads would then wait for the next
// k-tile's prefetch as well.
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(1))) u32x4 g_u32x4;
#else
typedef const u32x4 g_u32x4;    // host pass: never executed
#endif
__device__ inline u32x4 global_load16(const u32x4* p) { return *(g_u32x4*)p; }

// Agent-scope (cross-XCD coherent) 16-byte load and 8-byte store, for units
// that other workgroups of the same launch write.
__device__ inline u32x4 coherent_load16(const u32x4* p) {
  const unsigned long long* q = reinterpret_cast<const unsigned long long*>(p);
  const unsigned long long a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  u32x4 v;
  v.x = (unsigned)a;
  v.y = (unsigned)(a >> 32);
  v.z = (unsigned)b;
  v.w = (unsigned)(b >> 32);
  return v;
}
__device__ inline void coherent_store8(void* p, unsigned long long v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The split GEMM tile fed from pre-split operands (the accumulators, compute
// and store are the split tile's).
template <int NS, int BM, int BN, int WM, int WN>
struct PTile : mg_sg::Tile<NS, BM, BN, WM, WN> {
  typedef mg_sg::Tile<NS, BM, BN, WM, WN> Base;
  typedef typename Base::S S;
  static constexpr int NT = Base::NT;
  static constexpr int U = 2 * NS;                       // units per (row, group)
  static constexpr int AU = BM * (BK / 8) * U / NT;      // units per thread and k-tile, A
  static constexpr int BU = BN * (BK / 8) * U / NT;      // the same for B
  static_assert(AU * NT == BM * (BK / 8) * U && BU * NT == BN * (BK / 8) * U, "staging layout");

  // One k-tile of both operands in flight (registers).
  struct Stage {
    u32x4 a[AU], b[BU];
  };
  Stage cur;

  // The k-tile of groups [g0, g0 + 4) of A (rows at the tile origin, ldga
  // groups per row) and B (columns at the tile origin, ldgb groups per
  // column) into st; rows >= M and columns >= N read as 0.  A row's k-tile is
  // 4 U consecutive units, and consecutive threads take consecutive units, so
  // every load instruction reads contiguous memory across the wave.
  // COHERENT_B: B with agent-scope loads (B written by other workgroups of
  // this launch).  SNA: parts per real / imaginary part stored in A (at
  // least NS); its first NS parts of each are read (A split once into more
  // parts than a product uses).  PADDED: both operands hold whole tiles
  // (rows of A up to a multiple of BM, columns of B up to a multiple of BN,
  // zeros past M and N), so the loads need no bounds checks.
  template <bool COHERENT_B = false, int SNA = NS, bool PADDED = false>
  __device__ static void pfetch_to(Stage& st, const u32x4* __restrict__ A, int64_t ldga, const u32x4* B, int64_t ldgb,
                                   int64_t M, int64_t N, int64_t g0) {
    static_assert(SNA >= NS, "A holds fewer parts than the product uses");
    static_assert(!(PADDED && COHERENT_B), "padded fetches are plain loads");
    const int tid = Base::ltid();
#pragma unroll
    for (int c = 0; c < AU; ++c) {
      const int u = tid + c * NT, r = u / (4 * U), e = u - r * (4 * U);
      int64_t src;
      if constexpr (SNA == NS) {
        src = ((int64_t)r * ldga + g0) * U + e;
      } else {
        const int g = e / U, q = e - g * U;
        src = ((int64_t)r * ldga + g0 + g) * (2 * SNA) + (q / NS) * SNA + q % NS;
      }
      if constexpr (PADDED) st.a[c] = global_load16(A + src);
      else st.a[c] = r < M ? global_load16(A + src) : u32x4{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int c = 0; c < BU; ++c) {
      const int u = tid + c * NT, r = u / (4 * U), e = u - r * (4 * U);
      const u32x4* p = B + ((int64_t)r * ldgb + g0) * U + e;
      if constexpr (PADDED) {
        st.b[c] = global_load16(p);
      } else if (r < N) {
        if constexpr (COHERENT_B) st.b[c] = coherent_load16(p);
        else st.b[c] = global_load16(p);
      } else {
        st.b[c] = u32x4{0u, 0u, 0u, 0u};
      }
    }
  }
  template <bool COHERENT_B = false, int SNA = NS, bool PADDED = false>
  __device__ void pfetch(const u32x4* __restrict__ A, int64_t ldga, const u32x4* B, int64_t ldgb, int64_t M,
                         int64_t N, int64_t g0) {
    pfetch_to<COHERENT_B, SNA, PADDED>(cur, A, ldga, B, ldgb, M, N, g0);
  }

  // A fetched k-tile into the LDS planes (no conversion).
  __device__ static void pstash_from(const Stage& st, S& s) {
    const int tid = Base::ltid();
#pragma unroll
    for (int c = 0; c < AU; ++c) {
      const int u = tid + c * NT, r = u / (4 * U), e = u - r * (4 * U), g = e / U, q = e - g * U;
      *reinterpret_cast<u32x4*>(&s.a[q][r][g * 8]) = st.a[c];
    }
#pragma unroll
    for (int c = 0; c < BU; ++c) {
      const int u = tid + c * NT, r = u / (4 * U), e = u - r * (4 * U), g = e / U, q = e - g * U;
      *reinterpret_cast<u32x4*>(&s.b[q][r][g * 8]) = st.b[c];
    }
  }
  __device__ void pstash(S& s) const { pstash_from(cur, s); }

  // acc += A B over KG groups (a multiple of 4) from group 0 of A's rows and
  // B's columns (A holding SNA parts).
  template <bool COHERENT_B = false, int SNA = NS>
  __device__ void pmac(S& sm, const u32x4* __restrict__ A, int64_t ldga, const u32x4* B, int64_t ldgb, int64_t M,
                       int64_t N, int64_t KG) {
    const int64_t KT = KG / 4;
    pfetch<COHERENT_B, SNA>(A, ldga, B, ldgb, M, N, 0);
    for (int64_t kt = 0; kt < KT; ++kt) {
      __syncthreads();   // previous tile's reads are done
      pstash(sm);
      __syncthreads();
      if (kt + 1 < KT) pfetch<COHERENT_B, SNA>(A, ldga, B, ldgb, M, N, (kt + 1) * 4);
      this->compute(sm);
    }
  }

  // acc += A_0 B_0 + ... + A_{nsrc-1} B_{nsrc-1}, every source KG groups (a
  // multiple of 4) with ldg groups per row of A and per column of B
  // (PADDED: see pfetch_to).
  template <bool PADDED = false>
  __device__ void pmac_n(S& sm, const u32x4* const* A, const u32x4* const* B, int nsrc, int64_t ldg, int64_t M,
                         int64_t N, int64_t KG) {
    const int64_t KT = KG / 4, T = nsrc * KT;
    pfetch<false, NS, PADDED>(A[0], ldg, B[0], ldg, M, N, 0);
    for (int64_t t = 0; t < T; ++t) {
      __syncthreads();   // previous tile's reads are done
      pstash(sm);
      __syncthreads();
      if (t + 1 < T) {
        const int64_t src = (t + 1) / KT, kt = (t + 1) - src * KT;
        pfetch<false, NS, PADDED>(A[src], ldg, B[src], ldg, M, N, kt * 4);
      }
      this->compute(sm);
    }
  }

  // alpha * acc as split B-operand units of this tile's BM rows (k) and BN
  // columns: XS at the tile origin (row group 0, column 0), ldg groups per
  // column; rows >= M are written as 0, columns >= N not at all.  Agent-scope
  // stores (read by other workgroups of the launch).  Each lane holds 4
  // consecutive rows of a column: half a group, 8 bytes per plane.
  __device__ void store_split_cols(u32x4* XS, int64_t ldg, int64_t M, int64_t N, float2 alpha) const {
#pragma unroll
    for (int a = 0; a < Base::TM; ++a)
#pragma unroll
      for (int b = 0; b < Base::TN; ++b) {
        const int col = Base::col_of(b);
        if (col >= N) continue;
        const int row = Base::row_of(a, 0);   // multiple of 4
        unsigned short t[2 * NS][4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          float xr = 0.f, xi = 0.f;
          if (row + r < M) {
            const float ar = this->cr[a][b][r], ai = this->ci[a][b][r];
            xr = alpha.x * ar - alpha.y * ai;
            xi = alpha.x * ai + alpha.y * ar;
          }
          unsigned short p[NS];
          mg_sg::split_n<NS>(xr, p);
#pragma unroll
          for (int q = 0; q < NS; ++q) t[q][r] = p[q];
          mg_sg::split_n<NS>(xi, p);
#pragma unroll
          for (int q = 0; q < NS; ++q) t[NS + q][r] = p[q];
        }
        u32x4* unit = XS + ((int64_t)col * ldg + (row >> 3)) * U;
#pragma unroll
        for (int q = 0; q < U; ++q) {
          const unsigned long long v = (unsigned long long)(t[q][0] | ((unsigned)t[q][1] << 16)) |
                                       ((unsigned long long)(t[q][2] | ((unsigned)t[q][3] << 16)) << 32);
          coherent_store8(reinterpret_cast<unsigned long long*>(unit + q) + ((row >> 2) & 1), v);
        }
      }
  }
};

}  // namespace mg_ps
