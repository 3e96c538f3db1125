// Diagonal blocks of the inverse of a matrix from its pivoted LU factors.
//
// The LDOS at the scatterers needs diag(G W G^T) with W = M^{-1} and G the
// Green's matrix between the sites, which the system matrix itself holds
// (G = (I' - M) / (k0^2 alpha) off the self blocks): every term then reduces
// to the d x d diagonal blocks of W.  The reference forms all of W by
// solve(M, eye) (Transmission2D.py:416, Transmission3D.py:403), (32/3) n^3
// real flops; here W's diagonal blocks come from the factors P M = L U as
// W = U^{-1} L^{-1} P:
//   1. U^{-1} and L^{-1} overwrite U and L in place, block column by block
//      column (n^3 / 6 complex MACs each), every step a triangular product
//      (the GEMM body with a triangle-masked A operand, k-tiles outside the
//      triangle skipped) and a GEMM with the stored inverse of the diagonal
//      block; the two halves touch disjoint triangles and run on two streams;
//   2. W[r, c] = sum_k U^{-1}[r, k] L^{-1}[k, q(c)] for r, c in one site's
//      block (q: the column of L^{-1} P that holds column c, from ipiv),
//      a gather over one row of U^{-1} and one column of L^{-1} per entry.
// Total (16/3) n^3 real flops with the factorisation, against (32/3) n^3.
#include "cgemm.h"
#include "mg_internal.h"
#include "lu.h"

namespace {

using mg::cd;
using mg::kInvBlock;

// Y = T X with T a triangular principal submatrix of LU storage.
template <int BM, int BN, int BK, int WM, int WN, bool G3>
__global__ __launch_bounds__(64 * WM * WN) void k_trmm(int64_t M, int64_t N, int64_t K, TriA a, const void* B,
                                                       int64_t ldb, void* C, int64_t ldc, int nbm, int nbn) {
  int tm, tn;
  mg_tile_order(blockIdx.x, nbm, nbn, tm, tn);
  double2 one, zero;
  one.x = 1.0; one.y = 0.0;
  zero.x = 0.0; zero.y = 0.0;
  cgemm_body<double, BM, BN, BK, WM, WN, TriA, G3>(M, N, K, one, a, B, ldb, zero, C, ldc, tm, tn);
}

int trmm(int64_t M, int64_t N, int64_t K, const TriA& a, const void* B, int64_t ldb, void* C, int64_t ldc,
         hipStream_t s) {
  if (M <= 0 || N <= 0) return MG_OK;
  const int64_t big = ((M + 127) / 128) * ((N + 63) / 64);
  if (big >= (int64_t)mg::cu_count()) {
    const int nbm = (int)((M + 127) / 128), nbn = (int)((N + 63) / 64);
    hipLaunchKernelGGL((k_trmm<128, 64, 16, 4, 2, true>), dim3(nbm * nbn), dim3(512), 0, s, M, N, K, a, B, ldb, C,
                       ldc, nbm, nbn);
  } else {
    const int nbm = (int)((M + 63) / 64), nbn = (int)((N + 63) / 64);
    hipLaunchKernelGGL((k_trmm<64, 64, 16, 2, 2, false>), dim3(nbm * nbn), dim3(256), 0, s, M, N, K, a, B, ldb, C,
                       ldc, nbm, nbn);
  }
  MG_HIP(hipGetLastError());
  return MG_OK;
}

// The upper triangle with the diagonal (upper) or the strict lower triangle
// of the h x h block D (ld nb) into A (lda); the other triangle of A is kept.
__global__ __launch_bounds__(256) void k_copy_tri(const double2* __restrict__ D, int64_t nb, double2* A, int64_t lda,
                                                  int h, int upper) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)h * h) return;
  const int r = (int)(e / h), c = (int)(e % h);
  if (upper ? c >= r : c < r) A[r * lda + c] = D[r * nb + c];
}

int copy_tri(const double2* D, double2* A, int64_t lda, int h, bool upper, hipStream_t s) {
  const int64_t t = (int64_t)h * h;
  hipLaunchKernelGGL(k_copy_tri, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s, D, (int64_t)kInvBlock, A, lda,
                     h, upper ? 1 : 0);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

__device__ inline double2 cfma(double2 a, double2 b, double2 c) {
  c.x += a.x * b.x - a.y * b.y;
  c.y += a.x * b.y + a.y * b.x;
  return c;
}

// One workgroup per group of G sites: out[site][a][b] = W[d s + a, d s + b]
// = sum_k Uinv[d s + a, k] Linv[k, q(d s + b)] over the inverted LU storage
// (Uinv: upper triangle with the diagonal; Linv: strict lower triangle,
// unit diagonal).  Threads stride over k; the sums are reduced in a fixed
// order (lanes by a shuffle tree, then the waves in order).
template <int D, int G>
__global__ __launch_bounds__(256) void k_inv_diag_blocks(const double2* __restrict__ LU, int64_t n, int64_t lda,
                                                         const int* __restrict__ colperm,
                                                         const int64_t* __restrict__ sites, int64_t nsites,
                                                         double2* __restrict__ out) {
  constexpr int R = G * D;
  __shared__ double2 red[4][G * D * D];
  const int64_t s0 = (int64_t)blockIdx.x * G;
  int64_t row[R], col[R];
  bool ok[R];
  int64_t kmin = n;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int64_t s = s0 + i / D;
    // a site index outside [0, n / D) reads nothing and yields a NaN block
    ok[i] = s < nsites && sites[s] >= 0 && D * sites[s] + D <= n;
    row[i] = ok[i] ? D * sites[s] + i % D : 0;
    col[i] = ok[i] ? (int64_t)colperm[row[i]] : 0;
    if (ok[i]) {
      kmin = row[i] < kmin ? row[i] : kmin;
      kmin = col[i] < kmin ? col[i] : kmin;
    }
  }
  double2 acc[G][D][D];
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) acc[g][a][b] = double2{0.0, 0.0};
  for (int64_t k = kmin + threadIdx.x; k < n; k += 256) {
    double2 u[R], l[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      u[i] = (ok[i] && k >= row[i]) ? LU[row[i] * lda + k] : double2{0.0, 0.0};
      if (ok[i] && k > col[i]) l[i] = LU[k * lda + col[i]];
      else l[i] = double2{(ok[i] && k == col[i]) ? 1.0 : 0.0, 0.0};
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) acc[g][a][b] = cfma(u[g * D + a], l[g * D + b], acc[g][a][b]);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) {
        double2 v = acc[g][a][b];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          v.x += __shfl_xor(v.x, off);
          v.y += __shfl_xor(v.y, off);
        }
        if (lane == 0) red[wave][(g * D + a) * D + b] = v;
      }
  __syncthreads();
  for (int e = threadIdx.x; e < G * D * D; e += 256) {
    const int g = e / (D * D);
    if (s0 + g >= nsites) continue;
    double2 v = red[0][e];
#pragma unroll
    for (int w = 1; w < 4; ++w) {
      v.x += red[w][e].x;
      v.y += red[w][e].y;
    }
    const int64_t st = sites[s0 + g];
    if (st < 0 || D * st + D > n) v = double2{__builtin_nan(""), __builtin_nan("")};
    out[(s0 + g) * D * D + e % (D * D)] = v;
  }
}

template <int D, int G>
int gather(const double2* LU, int64_t n, int64_t lda, const int* colperm, const int64_t* sites, int64_t nsites,
           double2* out, hipStream_t s) {
  const int64_t groups = (nsites + G - 1) / G;
  hipLaunchKernelGGL((k_inv_diag_blocks<D, G>), dim3((unsigned)groups), dim3(256), 0, s, LU, n, lda, colperm, sites,
                     nsites, out);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

}  // namespace

namespace mg {

int64_t inverse_diag_work_elems(int64_t n) { return 2 * n * kInvBlock; }

int inverse_diag_blocks(Workspace& ws, void* LUp, int64_t n, int64_t lda, const void* dinv, const int* colperm,
                        int d, const int64_t* sites, int64_t nsites, void* outp, void* work, int64_t work_elems,
                        hipStream_t s) {
  if (n < 0 || lda < n || (d != 1 && d != 2 && d != 3) || nsites < 0) return MG_ERR_ARG;
  if (work_elems < inverse_diag_work_elems(n)) {
    set_error("inverse_diag_blocks: work needs %lld elements", (long long)inverse_diag_work_elems(n));
    return MG_ERR_ARG;
  }
  if (n == 0 || nsites == 0) return MG_OK;
  double2* LU = reinterpret_cast<double2*>(LUp);
  const double2* Dv = reinterpret_cast<const double2*>(dinv);
  double2* YU = reinterpret_cast<double2*>(work);
  double2* YL = YU + n * kInvBlock;
  const int64_t nb = kInvBlock, nblk = (n + nb - 1) / nb, bs = nb * nb;
  hipStream_t sl = ws.s2 ? ws.s2 : s;
  if (sl != s) {
    MG_HIP(hipEventRecord(ws.evA, s));
    MG_HIP(hipStreamWaitEvent(sl, ws.evA, 0));
  }
  // U^{-1}: block columns left to right; column J above the diagonal is
  // -U11^{-1} U12 inv(U_JJ) with U11^{-1} the part already inverted.
  for (int64_t J = 0; J < nblk; ++J) {
    const int64_t j = J * nb, jb = (n - j < nb) ? n - j : nb;
    if (j > 0) {
      MG_TRY(trmm(j, jb, j, TriA{LU, lda, TriA::UPPER}, LU + j, lda, YU, nb, s));
      MG_TRY(cgemm<double>(j, jb, jb, cd{-1, 0}, YU, nb, Dv + (nblk + J) * bs, nb, cd{0, 0}, LU + j, lda, s));
    }
    MG_TRY(copy_tri(Dv + (nblk + J) * bs, LU + j * lda + j, lda, (int)jb, true, s));
  }
  // L^{-1}: block columns right to left; column J below the diagonal is
  // -L22^{-1} L21 inv(L_JJ) with L22^{-1} the part already inverted.
  for (int64_t J = nblk - 1; J >= 0; --J) {
    const int64_t j = J * nb, jb = (n - j < nb) ? n - j : nb, r0 = j + jb, m = n - r0;
    if (m > 0) {
      MG_TRY(trmm(m, jb, m, TriA{LU + r0 * lda + r0, lda, TriA::LOWER_UNIT}, LU + r0 * lda + j, lda, YL, nb, sl));
      MG_TRY(cgemm<double>(m, jb, jb, cd{-1, 0}, YL, nb, Dv + J * bs, nb, cd{0, 0}, LU + r0 * lda + j, lda, sl));
    }
    MG_TRY(copy_tri(Dv + J * bs, LU + j * lda + j, lda, (int)jb, false, sl));
  }
  if (sl != s) {
    MG_HIP(hipEventRecord(ws.evB, sl));
    MG_HIP(hipStreamWaitEvent(s, ws.evB, 0));
  }
  double2* out = reinterpret_cast<double2*>(outp);
  if (d == 1) return gather<1, 32>(LU, n, lda, colperm, sites, nsites, out, s);
  if (d == 2) return gather<2, 8>(LU, n, lda, colperm, sites, nsites, out, s);
  return gather<3, 3>(LU, n, lda, colperm, sites, nsites, out, s);
}

}  // namespace mg
