// Blocked complex LU without pivoting, the factorisation behind the
// mixed-precision solve.  The factors only precondition an FP64 GMRES
// refinement, which restores full accuracy; a zero pivot sets info and the
// caller falls back to the pivoted factorisation.  Without pivoting there is
// no column search, no grid barrier and no row swap: each 256-column step is
// a small diagonal-block LU, two triangular solves (U12 = L11^{-1} A12,
// L21 = A21 U11^{-1}) and the trailing GEMM.
#include "mg_internal.h"
#include "lu.h"

namespace {

template <typename T> struct C2;
template <> struct C2<double> { typedef double2 t; };
template <> struct C2<float> { typedef float2 t; };

template <typename T2>
__device__ inline T2 crecip(T2 a) {
  // Smith's method for 1/(x + iy)
  T2 r;
  if (fabs(a.x) >= fabs(a.y)) {
    const auto q = a.y / a.x, d = a.x + a.y * q;
    r.x = 1 / d; r.y = -q / d;
  } else {
    const auto q = a.x / a.y, d = a.y + a.x * q;
    r.x = q / d; r.y = -1 / d;
  }
  return r;
}

template <typename T2>
__device__ inline T2 cmul(T2 a, T2 b) {
  T2 r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// LU without pivoting of an nb x nb block (nb <= 32) held in LDS by one
// workgroup; a zero pivot records its global index (1-based) in info and
// leaves the column unscaled.
template <typename T>
__global__ __launch_bounds__(256) void k_lu_leaf_nopiv(typename C2<T>::t* __restrict__ A, int64_t lda, int nb,
                                                       int64_t row0, int* __restrict__ info) {
  typedef typename C2<T>::t T2;
  __shared__ T2 S[mg::kLeaf][mg::kLeaf + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < nb * nb; e += blockDim.x) S[e / nb][e % nb] = A[(e / nb) * lda + e % nb];
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    const T2 piv = S[j][j];
    const bool nz = piv.x != 0 || piv.y != 0;
    if (!nz && tid == 0) atomicCAS(info, 0, (int)(row0 + j + 1));
    if (nz && tid > j && tid < nb) S[tid][j] = cmul(S[tid][j], crecip(piv));
    __syncthreads();
    const int w = nb - j - 1;
    for (int e = tid; e < w * w; e += blockDim.x) {
      const int r = j + 1 + e / w, c = j + 1 + e % w;
      const T2 l = S[r][j], u = S[j][c];
      S[r][c].x -= l.x * u.x - l.y * u.y;
      S[r][c].y -= l.x * u.y + l.y * u.x;
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * nb; e += blockDim.x) A[(e / nb) * lda + e % nb] = S[e / nb][e % nb];
}

// Right-side solve X U = B for an nb x nb (nb <= 32) upper triangular U with
// its diagonal, in place of B (nrows x nb).  Each half-wave holds one row of B
// with one column per lane: x_j = b_j / u_jj is broadcast by a lane shuffle
// and the later columns update in parallel.  8 rows per 256 threads.
template <typename T>
__global__ __launch_bounds__(256) void k_trsm_right_upper_leaf(const typename C2<T>::t* __restrict__ U, int64_t ldu,
                                                               int nb, typename C2<T>::t* __restrict__ B,
                                                               int64_t ldb, int64_t nrows) {
  typedef typename C2<T>::t T2;
  __shared__ T2 Us[mg::kLeaf][mg::kLeaf + 1];
  __shared__ T2 dinv[mg::kLeaf];
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) Us[e / nb][e % nb] = U[(e / nb) * ldu + e % nb];
  __syncthreads();
  if (threadIdx.x < nb) dinv[threadIdx.x] = crecip(Us[threadIdx.x][threadIdx.x]);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int c = lane & 31, half = lane & 32;
  const int64_t row = (int64_t)blockIdx.x * 8 + (threadIdx.x >> 5);
  const bool live = row < nrows && c < nb;
  T2 x;
  x.x = 0; x.y = 0;
  if (live) x = B[row * ldb + c];
  for (int j = 0; j < nb; ++j) {
    if (c == j) x = cmul(x, dinv[j]);
    T2 xj;
    xj.x = __shfl(x.x, half + j);
    xj.y = __shfl(x.y, half + j);
    if (c > j && c < nb) {
      const T2 u = Us[j][c];
      x.x -= xj.x * u.x - xj.y * u.y;
      x.y -= xj.x * u.y + xj.y * u.x;
    }
  }
  if (live) B[row * ldb + c] = x;
}

__global__ void k_zero_int(int* p) { *p = 0; }

inline int64_t split_half(int64_t w) {
  int64_t h = ((w / 2) + 15) / 16 * 16;
  if (h >= w) h = w / 2;
  return h > 0 ? h : 1;
}

}  // namespace

namespace mg {

// X U = B in place (B nrows x nb), U upper with its diagonal: recursion on nb.
template <typename T>
static int trsm_right_upper(const void* U, int64_t ldu, int64_t nb, void* B, int64_t ldb, int64_t nrows,
                            hipStream_t s) {
  typedef typename C2<T>::t T2;
  if (nb <= 0 || nrows <= 0) return MG_OK;
  if (nb <= kLeaf) {
    hipLaunchKernelGGL(k_trsm_right_upper_leaf<T>, dim3((unsigned)((nrows + 7) / 8)), dim3(256), 0, s,
                       reinterpret_cast<const T2*>(U), ldu, (int)nb, reinterpret_cast<T2*>(B), ldb, nrows);
    MG_HIP(hipGetLastError());
    return MG_OK;
  }
  const int64_t n1 = split_half(nb);
  const T2* Up = reinterpret_cast<const T2*>(U);
  T2* Bp = reinterpret_cast<T2*>(B);
  // [X1 X2] [[U11 U12] [0 U22]] = [B1 B2]: X1 = B1 U11^{-1}; B2 -= X1 U12; X2 = B2 U22^{-1}
  MG_TRY(trsm_right_upper<T>(Up, ldu, n1, Bp, ldb, nrows, s));
  MG_TRY(cgemm<T>(nrows, nb - n1, n1, cd{-1, 0}, Bp, ldb, Up + n1, ldu, cd{1, 0}, Bp + n1, ldb, s));
  return trsm_right_upper<T>(Up + n1 * ldu + n1, ldu, nb - n1, Bp + n1, ldb, nrows, s);
}

// Unpivoted LU of an h x h diagonal block in place (recursive).
template <typename T>
static int lu_block_nopiv(void* D, int64_t ld, int64_t h, int64_t row0, int* info, hipStream_t s) {
  typedef typename C2<T>::t T2;
  if (h <= kLeaf) {
    hipLaunchKernelGGL(k_lu_leaf_nopiv<T>, dim3(1), dim3(256), 0, s, reinterpret_cast<T2*>(D), ld, (int)h, row0,
                       info);
    MG_HIP(hipGetLastError());
    return MG_OK;
  }
  const int64_t h1 = split_half(h);
  T2* P = reinterpret_cast<T2*>(D);
  MG_TRY(lu_block_nopiv<T>(P, ld, h1, row0, info, s));
  MG_TRY(trsm_lower_unit<T>(P, ld, h1, P + h1, ld, h - h1, s));
  MG_TRY(trsm_right_upper<T>(P, ld, h1, P + h1 * ld, ld, h - h1, s));
  MG_TRY(cgemm<T>(h - h1, h - h1, h1, cd{-1, 0}, P + h1 * ld, ld, P + h1, ld, cd{1, 0}, P + h1 * ld + h1, ld, s));
  return lu_block_nopiv<T>(P + h1 * ld + h1, ld, h - h1, row0 + h1, info, s);
}

template <typename T>
int getrf_nopiv(Workspace& ws, void* A, int64_t n, int64_t lda, int* info, hipStream_t s, GemmFn update) {
  typedef typename C2<T>::t T2;
  (void)ws;
  if (n < 0 || lda < n) return MG_ERR_ARG;
  T2* Ap = reinterpret_cast<T2*>(A);
  const int64_t nb = kInvBlock;
  GemmFn fn = update ? update : &cgemm_trail<T>;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, info);
  for (int64_t k = 0; k < n; k += nb) {
    const int64_t h = (n - k < nb) ? n - k : nb, kn = k + h, r = n - kn;
    MG_TRY(lu_block_nopiv<T>(Ap + k * lda + k, lda, h, k, info, s));
    if (r > 0) {
      MG_TRY(trsm_lower_unit<T>(Ap + k * lda + k, lda, h, Ap + k * lda + kn, lda, r, s));
      MG_TRY(trsm_right_upper<T>(Ap + k * lda + k, lda, h, Ap + kn * lda + k, lda, r, s));
      MG_TRY(fn(r, r, h, cd{-1, 0}, Ap + kn * lda + k, lda, Ap + k * lda + kn, lda, cd{1, 0}, Ap + kn * lda + kn, lda,
                s));
    }
  }
  return MG_OK;
}

template int getrf_nopiv<double>(Workspace&, void*, int64_t, int64_t, int*, hipStream_t, GemmFn);
template int getrf_nopiv<float>(Workspace&, void*, int64_t, int64_t, int*, hipStream_t, GemmFn);

}  // namespace mg
