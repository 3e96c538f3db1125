// Complex GEMM on CDNA MFMA: C = alpha * A * B + beta * C (row-major,
// interleaved complex).  Complex products run as four real MFMA products
// (Cr += Ar Br - Ai Bi, Ci += Ar Bi + Ai Br) over LDS tiles that hold the
// real and imaginary planes separately.
//
// The A tile comes from a loader policy: MemA reads a row-major matrix,
// GreenA<KIND> evaluates the Green's function of (row point, column point)
// pairs in FP64 straight into LDS, so the propagation operator is never
// stored.
#pragma once
#include "mg_common.h"
#include "green.h"
#include <type_traits>

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Cplx;
template <> struct Cplx<double> { typedef double2 t; typedef v4d acc; };
template <> struct Cplx<float> { typedef float2 t; typedef v4f acc; };

__device__ inline v4d mg_mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline v4f mg_mfma(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// c - a b: on gfx950 the F64 MFMA reads its blgp field as per-operand negate
// bits (blgp = 1 assembles as neg:[1,0,0]), so the sign flip costs nothing;
// the F32 MFMA has no such modifier.
__device__ inline v4d mg_mfma_nega(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 1);
}
__device__ inline v4f mg_mfma_nega(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(-a, b, c, 0, 0, 0);
}

// Map a linear workgroup id to (tile_m, tile_n) so that the tiles one XCD
// works on are contiguous groups of rows: ids are dealt round-robin over the
// 8 XCDs, so undo that first (for the leading multiple of 8 ids; the few
// left over keep their place), then walk the tile grid in row groups of 8.
// When an XCD's share is a whole number of tile rows, the XCD takes exactly
// those rows (every tile of a row block on one L2: a 12288 x 256 product's
// 96 x 4 tiles read M once from HBM instead of twice).
__device__ inline void mg_tile_order(int bid, int nbm, int nbn, int& tm, int& tn) {
  const int total = nbm * nbn;
  const int per = total / 8;
  int id = bid, row0 = 0, rows = nbm;
  if (total >= 64 && bid < 8 * per) {
    if (per % nbn == 0) {
      rows = per / nbn;
      row0 = (bid % 8) * rows;
      id = bid / 8;
    } else {
      id = (bid % 8) * per + bid / 8;
    }
  }
  const int group = 8;
  const int in_group = group * nbn;
  const int g = id / in_group;
  const int first_m = g * group;
  const int gsize = min(rows - first_m, group);
  tm = row0 + first_m + (id % in_group) % gsize;
  tn = (id % in_group) / gsize;
}

struct MemA {
  const void* A;
  int64_t lda;
};

// Matrix elements read and written through the global address space: the
// operands arrive as untyped pointers, and flat accesses would also count
// against the LDS counter, so the waits for a k-tile's LDS reads would wait
// for the next k-tile's prefetch too.
// (The qualifier applies to the device pass; the host pass never runs these
// bodies and its vector types take no address-space-qualified operands.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
using gcptr = const __attribute__((address_space(1))) T*;
#else
template <typename T>
using gptr = T*;
template <typename T>
using gcptr = const T*;
#endif

// A triangular operand read from LU-factor storage (both triangles share
// one matrix): UPPER takes the upper triangle with the diagonal (entries
// below it read as 0); LOWER_UNIT the strict lower triangle with a unit
// diagonal (entries above it read as 0).  Row and column indices are those
// of the operand, a principal submatrix of the storage.
struct TriA {
  enum { UPPER = 0, LOWER_UNIT = 1 };
  const void* A;
  int64_t lda;
  int mode;
};

template <int KIND>
struct GreenA {
  static constexpr int kind = KIND;
  const double* rows;   // row points (space dim S each)
  const double* cols;   // column points
  int64_t nrows_pts, ncols_pts;
  double k0;
  int flags;
  double radius;
};

template <typename X>
struct IsGreenA : std::false_type {};
template <int KIND>
struct IsGreenA<GreenA<KIND>> : std::true_type {};

// LDS planes for one stage: [re, im (, re + im)][BK][BM + PAD].  The third
// plane serves the three-product (Gauss) complex multiply.
template <typename T, int BM, int BN, int BK, int NP = 2>
struct GemmSmem {
  T a[NP][BK][BM + 1];
  T b[NP][BK][BN + 1];
};

template <typename T, int BM, int BK, int NT>
struct ARegs {
  static constexpr int PER = (BM * BK + NT - 1) / NT;
  typename Cplx<T>::t v[PER];
};

// ---- A-tile staging: memory ----------------------------------------------
template <typename T, int BM, int BK, int NT>
__device__ inline void a_fetch(const MemA& src, int64_t M, int64_t K, int64_t i0, int64_t k0,
                               ARegs<T, BM, BK, NT>& r) {
  typedef typename Cplx<T>::t T2;
  gcptr<T2> A = (gcptr<T2>)src.A;
#pragma unroll
  for (int p = 0; p < r.PER; ++p) {
    const int e = p * NT + threadIdx.x;
    const int row = e / BK, kk = e % BK;
    T2 v;
    v.x = 0; v.y = 0;
    if (e < BM * BK && i0 + row < M && k0 + kk < K) v = A[(i0 + row) * src.lda + k0 + kk];
    r.v[p] = v;
  }
}

template <typename T, int BM, int BK, int NT>
__device__ inline void a_fetch(const TriA& src, int64_t M, int64_t K, int64_t i0, int64_t k0,
                               ARegs<T, BM, BK, NT>& r) {
  typedef typename Cplx<T>::t T2;
  gcptr<T2> A = (gcptr<T2>)src.A;
  const bool upper = src.mode == TriA::UPPER;
#pragma unroll
  for (int p = 0; p < r.PER; ++p) {
    const int e = p * NT + threadIdx.x;
    const int64_t row = i0 + e / BK, k = k0 + e % BK;
    T2 v;
    v.x = 0; v.y = 0;
    if (e < BM * BK && row < M && k < K) {
      if (upper ? k >= row : k < row) v = A[row * src.lda + k];
      else if (k == row) v.x = 1;
    }
    r.v[p] = v;
  }
}

// k-range of the product that can hold nonzero A entries for rows [i0, i0 + BM).
__device__ inline void a_krange(const MemA&, int64_t, int, int64_t K, int64_t& lo, int64_t& hi) { lo = 0; hi = K; }
template <int KIND>
__device__ inline void a_krange(const GreenA<KIND>&, int64_t, int, int64_t K, int64_t& lo, int64_t& hi) {
  lo = 0;
  hi = K;
}
__device__ inline void a_krange(const TriA& src, int64_t i0, int BM, int64_t K, int64_t& lo, int64_t& hi) {
  if (src.mode == TriA::UPPER) {
    lo = i0 < K ? i0 : K;
    hi = K;
  } else {
    lo = 0;
    hi = i0 + BM < K ? i0 + BM : K;
  }
}

template <typename T, int BM, int BN, int BK, int NT, int NP>
__device__ inline void a_store(GemmSmem<T, BM, BN, BK, NP>& s, const ARegs<T, BM, BK, NT>& r) {
#pragma unroll
  for (int p = 0; p < r.PER; ++p) {
    const int e = p * NT + threadIdx.x;
    if (e < BM * BK) {
      const int row = e / BK, kk = e % BK;
      s.a[0][kk][row] = r.v[p].x;
      s.a[1][kk][row] = r.v[p].y;
      if constexpr (NP == 3) s.a[2][kk][row] = r.v[p].x + r.v[p].y;
    }
  }
}

// ---- A-tile staging: Green's function evaluated in FP64 ------------------
// Tile rows i0..i0+BM are d*point + a, tile columns k0..k0+BK are
// d*scatterer + b; BM and BK are multiples of d.
template <int KIND, typename T, int BM, int BN, int BK, int NT, int NP>
__device__ inline void a_green(GemmSmem<T, BM, BN, BK, NP>& s, const GreenA<KIND>& src, int64_t i0, int64_t k0) {
  constexpr int D = (KIND == MG_2D_TE) ? 2 : (KIND == MG_3D_VECTOR ? 3 : 1);
  constexpr int SP = (KIND == MG_2D_TM || KIND == MG_2D_TE) ? 2 : 3;
  constexpr int PM = BM / D, PK = BK / D;
  static_assert(BM % D == 0 && BK % D == 0, "tile must hold whole blocks");
  const int64_t p0 = i0 / D, q0 = k0 / D;
  for (int e = threadIdx.x; e < PM * PK; e += NT) {
    const int pm = e / PK, pk = e % PK;
    const int64_t pi = p0 + pm, qj = q0 + pk;
    mg_c g[D * D];
    if (pi < src.nrows_pts && qj < src.ncols_pts) {
      mg_green<KIND>(src.rows + pi * SP, src.cols + qj * SP, src.k0, src.flags, src.radius, g);
    } else {
#pragma unroll
      for (int t = 0; t < D * D; ++t) g[t] = {0.0, 0.0};
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
#pragma unroll
      for (int b = 0; b < D; ++b) {
        s.a[0][pk * D + b][pm * D + a] = (T)g[a * D + b].re;
        s.a[1][pk * D + b][pm * D + a] = (T)g[a * D + b].im;
        if constexpr (NP == 3) s.a[2][pk * D + b][pm * D + a] = (T)(g[a * D + b].re + g[a * D + b].im);
      }
    }
  }
}

// ---- B tile ----------------------------------------------------------------
template <typename T, int BN, int BK, int NT>
struct BRegs {
  static constexpr int PER = (BN * BK + NT - 1) / NT;
  typename Cplx<T>::t v[PER];
};

template <typename T, int BN, int BK, int NT>
__device__ inline void b_fetch(const void* Bp, int64_t ldb, int64_t K, int64_t N, int64_t k0, int64_t j0,
                               BRegs<T, BN, BK, NT>& r) {
  typedef typename Cplx<T>::t T2;
  gcptr<T2> B = (gcptr<T2>)Bp;
#pragma unroll
  for (int p = 0; p < r.PER; ++p) {
    const int e = p * NT + threadIdx.x;
    const int kk = e / BN, col = e % BN;
    T2 v;
    v.x = 0; v.y = 0;
    if (e < BN * BK && k0 + kk < K && j0 + col < N) v = B[(k0 + kk) * ldb + j0 + col];
    r.v[p] = v;
  }
}

template <typename T, int BM, int BN, int BK, int NT, int NP>
__device__ inline void b_store(GemmSmem<T, BM, BN, BK, NP>& s, const BRegs<T, BN, BK, NT>& r) {
#pragma unroll
  for (int p = 0; p < r.PER; ++p) {
    const int e = p * NT + threadIdx.x;
    if (e < BN * BK) {
      const int kk = e / BN, col = e % BN;
      s.b[0][kk][col] = r.v[p].x;
      s.b[1][kk][col] = r.v[p].y;
      if constexpr (NP == 3) s.b[2][kk][col] = r.v[p].x + r.v[p].y;
    }
  }
}

// ---- the kernel body -------------------------------------------------------
// WM x WN waves of 64 lanes; each wave owns (BM/WM) x (BN/WN) of C as
// 16x16 MFMA tiles.  G3: complex products as three real MFMA products
// (Gauss: t1 = Ar Br, t2 = Ai Bi, t3 = (Ar + Ai)(Br + Bi); Cr = t1 - t2,
// Ci = t3 - t1 - t2) instead of four, at the cost of a third accumulator.
template <typename T, int BM, int BN, int BK, int WM, int WN, typename ASrc, bool G3 = false>
__device__ __attribute__((always_inline)) inline void cgemm_body(int64_t M, int64_t N, int64_t K, typename Cplx<T>::t alpha, const ASrc& asrc,
                                  const void* Bp, int64_t ldb, typename Cplx<T>::t beta, void* Cp, int64_t ldc,
                                  int tm, int tn) {
  typedef typename Cplx<T>::t T2;
  typedef typename Cplx<T>::acc ACC;
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
  constexpr int NP = G3 ? 3 : 2;
  typedef GemmSmem<T, BM, BN, BK, NP> Smem;
  static_assert(BM % (WM * 16) == 0 && BN % (WN * 16) == 0 && BK % 4 == 0, "tile shape");
  __shared__ Smem sm[2];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int64_t i0 = (int64_t)tm * BM, j0 = (int64_t)tn * BN;
  const int li = lane & 15, lk = lane >> 4;

  // k-tiles [kt0, kt1): those where A can be nonzero for this tile's rows
  int64_t klo, khi;
  a_krange(asrc, i0, BM, K, klo, khi);
  const int64_t kt0 = klo / BK, kt1 = (khi + BK - 1) / BK;
  constexpr bool kMem = !IsGreenA<ASrc>::value;    // A read from memory (MemA, TriA)
  ARegs<T, BM, BK, NT> ar;
  BRegs<T, BN, BK, NT> br;

  // prologue
  if constexpr (kMem) {
    a_fetch<T, BM, BK, NT>(asrc, M, K, i0, kt0 * BK, ar);
    a_store<T, BM, BN, BK, NT, NP>(sm[0], ar);
  } else {
    a_green<ASrc::kind, T, BM, BN, BK, NT, NP>(sm[0], asrc, i0, 0);
  }
  b_fetch<T, BN, BK, NT>(Bp, ldb, K, N, kt0 * BK, j0, br);

  // Accumulator item r of lane l holds column l%16 and row l/16 + 4r (f64
  // MFMA) or 4(l/16) + r (f32 MFMA), as measured on gfx950.  The
  // accumulators start from (beta / alpha) C, loaded here so that its
  // latency hides behind the first k-steps; the epilogue then only scales by
  // alpha and writes (alpha = 0 keeps C = beta C).
  // 4M: acc[0] = Cr, acc[1] = Ci.  3M: acc[0..2] = t1, t2, t3, started at
  // t1 = cr, t2 = 0, t3 = cr + ci so that t1 - t2 and t3 - t1 - t2 carry c.
  gptr<T2> C = (gptr<T2>)Cp;
  constexpr bool kRowsStrided = std::is_same<T, double>::value;
  const bool alpha0 = (alpha.x == 0 && alpha.y == 0);
  const bool cload = !(beta.x == 0 && beta.y == 0) && !alpha0;
  T2 q;
  q.x = 0;
  q.y = 0;
  if (cload) {
    const T d = alpha.x * alpha.x + alpha.y * alpha.y;
    q.x = (beta.x * alpha.x + beta.y * alpha.y) / d;
    q.y = (beta.y * alpha.x - beta.x * alpha.y) / d;
  }
  ACC acc[NP == 3 ? 3 : 2][TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a) {
#pragma unroll
    for (int b = 0; b < TN; ++b) {
      const int64_t col = j0 + wn * (BN / WN) + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = i0 + wm * (BM / WM) + a * 16 + (kRowsStrided ? lk + 4 * r : lk * 4 + r);
        T2 c;
        c.x = 0;
        c.y = 0;
        if (cload && row < M && col < N) c = C[row * ldc + col];
        const T cr = q.x * c.x - q.y * c.y, ci = q.x * c.y + q.y * c.x;
        acc[0][a][b][r] = cr;
        if constexpr (NP == 3) {
          acc[1][a][b][r] = 0;
          acc[2][a][b][r] = cr + ci;
        } else {
          acc[1][a][b][r] = ci;
        }
      }
    }
  }

  b_store<T, BM, BN, BK, NT, NP>(sm[0], br);
  __syncthreads();

  for (int64_t kt = kt0; kt < kt1; ++kt) {
    const int cur = (kt - kt0) & 1;
    const bool more = kt + 1 < kt1;
    if (more) {
      if constexpr (kMem) a_fetch<T, BM, BK, NT>(asrc, M, K, i0, (kt + 1) * BK, ar);
      b_fetch<T, BN, BK, NT>(Bp, ldb, K, N, (kt + 1) * BK, j0, br);
    }
    Smem& s = sm[cur];
#pragma unroll
    for (int g = 0; g < BK / 4; ++g) {
      T av[NP][TM], bv[NP][TN];
#pragma unroll
      for (int q = 0; q < NP; ++q) {
#pragma unroll
        for (int a = 0; a < TM; ++a) av[q][a] = s.a[q][g * 4 + lk][wm * (BM / WM) + a * 16 + li];
#pragma unroll
        for (int b = 0; b < TN; ++b) bv[q][b] = s.b[q][g * 4 + lk][wn * (BN / WN) + b * 16 + li];
      }
#pragma unroll
      for (int a = 0; a < TM; ++a) {
#pragma unroll
        for (int b = 0; b < TN; ++b) {
          if constexpr (NP == 3) {
            acc[0][a][b] = mg_mfma(av[0][a], bv[0][b], acc[0][a][b]);
            acc[1][a][b] = mg_mfma(av[1][a], bv[1][b], acc[1][a][b]);
            acc[2][a][b] = mg_mfma(av[2][a], bv[2][b], acc[2][a][b]);
          } else {
            acc[0][a][b] = mg_mfma(av[0][a], bv[0][b], acc[0][a][b]);
            acc[0][a][b] = mg_mfma_nega(av[1][a], bv[1][b], acc[0][a][b]);
            acc[1][a][b] = mg_mfma(av[0][a], bv[1][b], acc[1][a][b]);
            acc[1][a][b] = mg_mfma(av[1][a], bv[0][b], acc[1][a][b]);
          }
        }
      }
    }
    if (more) {
      Smem& nx = sm[cur ^ 1];
      if constexpr (kMem) a_store<T, BM, BN, BK, NT, NP>(nx, ar);
      else a_green<ASrc::kind, T, BM, BN, BK, NT, NP>(nx, asrc, i0, (kt + 1) * BK);
      b_store<T, BM, BN, BK, NT, NP>(nx, br);
    }
    __syncthreads();
  }

  // epilogue: C = alpha * acc (acc already holds (beta / alpha) C).
#pragma unroll
  for (int a = 0; a < TM; ++a) {
#pragma unroll
    for (int b = 0; b < TN; ++b) {
      const int64_t col = j0 + wn * (BN / WN) + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = i0 + wm * (BM / WM) + a * 16 + (kRowsStrided ? lk + 4 * r : lk * 4 + r);
        if (row < M && col < N) {
          T xr, xi;
          if constexpr (NP == 3) {
            const T t1 = acc[0][a][b][r], t2 = acc[1][a][b][r], t3 = acc[2][a][b][r];
            xr = t1 - t2;
            xi = t3 - t1 - t2;
          } else {
            xr = acc[0][a][b][r];
            xi = acc[1][a][b][r];
          }
          T2 out;
          out.x = alpha.x * xr - alpha.y * xi;
          out.y = alpha.x * xi + alpha.y * xr;
          if (alpha0) {
            const T2 c = C[row * ldc + col];
            out.x = beta.x * c.x - beta.y * c.y;
            out.y = beta.x * c.y + beta.y * c.x;
          }
          C[row * ldc + col] = out;
        }
      }
    }
  }
}
