// Relaxed-accuracy Krylov products y = A z for the FP64 refinement.
//
// A flexible GMRES whose residual has dropped to r can take its next product
// with a relative error of about tol / r and still end with a residual gap
// (true residual minus the Arnoldi estimate) of about tol: the error enters
// the final residual multiplied by that direction's coefficient, which is of
// the size of r (inexact Krylov methods; Simoncini and Szyld, SIAM J. Sci.
// Comput. 25 (2003); Bouras and Fraysse, SIAM J. Matrix Anal. Appl. 26
// (2005)).  The refinement therefore takes its last products as split-bf16
// MFMA products instead of FP64 ones:
//   A (complex128) is split once into three bf16 parts per real and
//   imaginary part (the pre-split row format of psplit.h, 12 bytes per
//   element); each product splits z into its own number of parts and runs
//   BF16x3 (first two parts of A, three products per real product, ~16 bits)
//   or BF16x6 (all three, six products, ~24 bits), split over K into FP32
//   partial sums that one fixed-order pass adds in FP64.
#include "cgemm.h"
#include "mg_internal.h"
#include "psplit.h"

namespace {

constexpr int kSrcNS = 3;   // parts of the stored split of A

int64_t groups_of(int64_t k) { return ((k + 7) / 8 + 3) / 4 * 4; }

// A (rows x cols, complex double or float, lda) -> three-part split rows, ldg
// groups per row.  A workgroup takes 2048 columns of one row: coalesced
// element loads into LDS, then each thread splits 8 consecutive elements and
// writes their 6 units (consecutive threads, consecutive units).
constexpr int kSplitSeg = 2048;

__device__ inline float2 to_f2(double2 v) { return make_float2((float)v.x, (float)v.y); }
__device__ inline float2 to_f2(float2 v) { return v; }

template <typename T2>
__global__ __launch_bounds__(256) void k_split_rows(const T2* __restrict__ A, int64_t lda, int64_t cols, int64_t ldg,
                                                    int segs, u32x4* __restrict__ AS) {
  __shared__ float2 buf[kSplitSeg + kSplitSeg / 8];   // a pad element after every 8
  const int64_t r = blockIdx.x / segs;
  const int64_t c0 = (int64_t)(blockIdx.x - r * segs) * kSplitSeg;
  const T2* row = A + r * lda;
#pragma unroll
  for (int p = 0; p < kSplitSeg / 256; ++p) {
    const int e = p * 256 + (int)threadIdx.x;
    const int64_t c = c0 + e;
    buf[e + e / 8] = c < cols ? to_f2(row[c]) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const int64_t g = c0 / 8 + threadIdx.x;
  if (g >= ldg) return;
  float2 v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = buf[threadIdx.x * 9 + e];
  mg_ps::split8<kSrcNS>(v, AS + (r * ldg + g) * (2 * kSrcNS));
}

// Z (K x N, complex128, ldz) -> NS-part split columns, ldg groups per column;
// one thread per (group, column), columns fastest (coalesced reads).
template <int NS>
__global__ void k_split_cols(const double2* __restrict__ Z, int64_t ldz, int64_t K, int64_t N, int64_t ldg,
                             u32x4* __restrict__ ZS) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * ldg) return;
  const int64_t g = t / N, c = t - g * N;
  float2 v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int64_t k = g * 8 + e;
    v[e] = make_float2(0.f, 0.f);
    if (k < K) {
      const double2 z = Z[k * ldz + c];
      v[e] = make_float2((float)z.x, (float)z.y);
    }
  }
  mg_ps::split8<NS>(v, ZS + (c * ldg + g) * (2 * NS));
}

// Partial products: workgroup (tile, part) forms A[tile rows, part's groups]
// B[part's groups, tile columns] into P + part M N (complex float, M x N
// row-major).  Tiles and parts are laid out as one grid of nbm x (nbn parts)
// tiles in the XCD-aware order, so that the workgroups reading one block of
// A rows share an XCD.
template <int NS, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void k_relaxed_product(const u32x4* __restrict__ AS, int64_t ldga,
                                                                  const u32x4* __restrict__ BS, int64_t ldgb,
                                                                  float2* __restrict__ P, int64_t M, int64_t N,
                                                                  int64_t KG, int64_t kgs, int nbm, int nbn,
                                                                  int parts) {
  typedef mg_ps::PTile<NS, BM, BN, WM, WN> Tile;
  __shared__ typename Tile::S sm;
  int tm, tq;
  mg_tile_order(blockIdx.x, nbm, nbn * parts, tm, tq);
  const int part = tq / nbn, tn = tq - part * nbn;
  const int64_t i0 = (int64_t)tm * BM, j0 = (int64_t)tn * BN;
  const int64_t g0 = part * kgs, ng = KG - g0 < kgs ? KG - g0 : kgs;
  Tile t;
  t.init(nullptr, 0, 0, 0, make_float2(0.f, 0.f), false);
  t.template pmac<false, kSrcNS>(sm, AS + i0 * ldga * (2 * kSrcNS) + g0 * (2 * kSrcNS), ldga,
                                 BS + j0 * ldgb * (2 * NS) + g0 * (2 * NS), ldgb, M - i0, N - j0, ng);
  t.store(P + part * M * N + i0 * N + j0, N, M - i0, N - j0, make_float2(1.f, 0.f), make_float2(0.f, 0.f), false);
}

// C = sum over parts of P (FP64 sums in part order).
__global__ void k_reduce_parts(const float2* __restrict__ P, int parts, int64_t M, int64_t N, double2* __restrict__ C,
                               int64_t ldc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M * N) return;
  const int64_t i = t / N, j = t - i * N;
  double re = 0.0, im = 0.0;
  for (int p = 0; p < parts; ++p) {
    const float2 v = P[p * M * N + t];
    re += (double)v.x;
    im += (double)v.y;
  }
  C[i * ldc + j] = make_double2(re, im);
}

template <int NS>
struct Shape;
template <>
struct Shape<2> {   // 64 x 128 tiles, 62 KiB of LDS: two workgroups per CU
  static constexpr int BM = 64, BN = 128, WM = 2, WN = 2;
};
template <>
struct Shape<3> {   // 64 x 64 tiles, 62 KiB of LDS: two workgroups per CU
  static constexpr int BM = 64, BN = 64, WM = 2, WN = 2;
};

// K parts: the fewest that give every CU two workgroups, preferring a grid
// that is a whole number of such rounds; every part a multiple of 4 groups.
int relaxed_parts(int64_t tiles, int64_t KG) {
  const int64_t slots = 2 * (int64_t)mg::cu_count();
  const int64_t maxp = std::max<int64_t>(1, std::min<int64_t>(16, KG / 64));
  int64_t best = 1;
  for (int64_t p = 1; p <= maxp; ++p) {
    best = p;
    if (tiles * p >= slots && (tiles * p) % slots == 0) break;
    if (tiles * p >= 4 * slots) break;
  }
  return (int)best;
}

template <int NS>
struct Plan {
  int nbm, nbn, parts;
  int64_t KG, kgs, pbytes, zbytes;
  Plan(int64_t M, int64_t N, int64_t K) {
    typedef Shape<NS> Sh;
    nbm = (int)((M + Sh::BM - 1) / Sh::BM);
    nbn = (int)((N + Sh::BN - 1) / Sh::BN);
    KG = groups_of(K);
    parts = relaxed_parts((int64_t)nbm * nbn, KG);
    kgs = ((KG + parts - 1) / parts + 3) / 4 * 4;
    parts = (int)((KG + kgs - 1) / kgs);
    pbytes = ((int64_t)parts * M * N * 8 + 255) / 256 * 256;
    zbytes = N * KG * 2 * NS * 16;
  }
  int64_t bytes() const { return pbytes + zbytes; }
};

template <int NS>
int relaxed_product(int64_t M, int64_t N, int64_t K, const void* AS, const void* B, int64_t ldb, void* C, int64_t ldc,
                    void* work, int64_t work_bytes, hipStream_t s) {
  typedef Shape<NS> Sh;
  const Plan<NS> pl(M, N, K);
  if (work == nullptr || work_bytes < pl.bytes()) {
    mg::set_error("relaxed product: work needs %lld bytes", (long long)pl.bytes());
    return MG_ERR_ARG;
  }
  float2* P = reinterpret_cast<float2*>(work);
  u32x4* ZS = reinterpret_cast<u32x4*>(reinterpret_cast<char*>(work) + pl.pbytes);
  const int64_t zt = N * pl.KG;
  hipLaunchKernelGGL((k_split_cols<NS>), dim3((unsigned)((zt + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(B), ldb, K, N, pl.KG, ZS);
  MG_HIP(hipGetLastError());
  const unsigned grid = (unsigned)((int64_t)pl.nbm * pl.nbn * pl.parts);
  hipLaunchKernelGGL((k_relaxed_product<NS, Sh::BM, Sh::BN, Sh::WM, Sh::WN>), dim3(grid), dim3(64 * Sh::WM * Sh::WN),
                     0, s, reinterpret_cast<const u32x4*>(AS), pl.KG, ZS, pl.KG, P, M, N, pl.KG, pl.kgs, pl.nbm,
                     pl.nbn, pl.parts);
  MG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)((M * N + 255) / 256)), dim3(256), 0, s, P, pl.parts, M, N,
                     reinterpret_cast<double2*>(C), ldc);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

}  // namespace

extern "C" {

int64_t mg_split_rows_bytes(int64_t rows, int64_t cols) {
  if (rows < 0 || cols < 0) return -1;
  return rows * groups_of(cols) * 2 * kSrcNS * 16;
}

int mg_split_rows(int prec, const void* A, int64_t lda, int64_t rows, int64_t cols, void* AS, void* stream) {
  if (rows < 0 || cols < 0 || lda < cols || (prec != MG_C128 && prec != MG_C64)) return MG_ERR_ARG;
  if (rows == 0 || cols == 0) return MG_OK;
  if (A == nullptr || AS == nullptr) return MG_ERR_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int64_t ldg = groups_of(cols);
  const int segs = (int)((ldg * 8 + kSplitSeg - 1) / kSplitSeg);
  const dim3 grid((unsigned)(rows * segs));
  u32x4* out = reinterpret_cast<u32x4*>(AS);
  if (prec == MG_C128)
    hipLaunchKernelGGL(k_split_rows<double2>, grid, dim3(256), 0, s, reinterpret_cast<const double2*>(A), lda, cols,
                       ldg, segs, out);
  else
    hipLaunchKernelGGL(k_split_rows<float2>, grid, dim3(256), 0, s, reinterpret_cast<const float2*>(A), lda, cols,
                       ldg, segs, out);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

int64_t mg_relaxed_product_work_bytes(int prec, int64_t m, int64_t n, int64_t k) {
  if (m <= 0 || n <= 0 || k <= 0) return 0;
  switch (prec) {
    case MG_BF16X3: return Plan<2>(m, n, k).bytes();
    case MG_BF16X6: return Plan<3>(m, n, k).bytes();
    default: return -1;
  }
}

int mg_relaxed_product(int prec, int64_t m, int64_t n, int64_t k, const void* AS, const void* B, int64_t ldb,
                       void* C, int64_t ldc, void* work, int64_t work_bytes, void* stream) {
  if (m < 0 || n < 0 || k < 0 || ldb < n || ldc < n) return MG_ERR_ARG;
  if (m == 0 || n == 0) return MG_OK;
  if (k == 0 || AS == nullptr || B == nullptr || C == nullptr) return MG_ERR_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  switch (prec) {
    case MG_BF16X3: return relaxed_product<2>(m, n, k, AS, B, ldb, C, ldc, work, work_bytes, s);
    case MG_BF16X6: return relaxed_product<3>(m, n, k, AS, B, ldb, C, ldc, work, work_bytes, s);
    default: return MG_ERR_ARG;
  }
}

}  // extern "C"
