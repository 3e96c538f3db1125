// Blocked complex LU with partial pivoting (getrf) and the multi-RHS solve
// (getrs), row-major, for complex double and complex float.
//
// getrf is right-looking with an outer block of NB columns.  Each NB-wide
// panel is factored recursively (left half, swap + triangular solve + GEMM
// on the right half, right half) down to leaves of at most 32 columns; a
// leaf is factored by k_getf2, a cooperative kernel whose workgroups each
// keep a slice of the leaf's rows in LDS and meet at one grid barrier per
// column.  The trailing update is the MFMA complex GEMM.  Pivoting follows
// LAPACK: the pivot of column j is the row of largest |re| + |im| on or
// below the diagonal, lowest row on ties; info is the first zero pivot
// (1-based); the factors overwrite A with unit-lower L and upper U, and
// ipiv[k] is the (0-based) row swapped with row k.
#include "mg_internal.h"
#include "lu.h"
#include <climits>

namespace {

constexpr int kThreads = 256;
constexpr int kBarWords = 1024;   // one set: top counter + 8 group counters, 64 words apart, failure word
constexpr int kErrWord = 1000;    // the set's failure word (past the group counters at 64 .. 512)
constexpr long long kSpinLimit = 1ll << 22;   // polls (~1 s) before a barrier gives up

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

__device__ inline bool better(double v, int r, double bv, int br) { return v > bv || (v == bv && r < br); }

// (value, row) arg-max over the workgroup; result valid in every thread.
__device__ inline void block_argmax(double& v, int& r, double* sv, int* sr) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double ov = __shfl_xor(v, off);
    const int orow = __shfl_xor(r, off);
    if (better(ov, orow, v, r)) { v = ov; r = orow; }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) { sv[wave] = v; sr[wave] = r; }
  __syncthreads();
  v = sv[0]; r = sr[0];
#pragma unroll
  for (int w = 1; w < kThreads / 64; ++w)
    if (better(sv[w], sr[w], v, r)) { v = sv[w]; r = sr[w]; }
  __syncthreads();
}

// Data exchanged between the workgroups of the leaf goes through agent-scope
// atomic stores and loads, which are coherent across XCDs without the L2
// write-back / invalidate that agent-scope fences cost (measured on MI355X:
// ~3 us vs ~31 us per barrier at 192 workgroups).
template <typename V>
__device__ inline void pub_store(V* p, V v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename V>
__device__ inline V pub_load(const V* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename T2>
__device__ inline void pub_store2(T2* p, T2 v) { pub_store(&p->x, v.x); pub_store(&p->y, v.y); }
template <typename T2>
__device__ inline T2 pub_load2(const T2* p) { T2 v; v.x = pub_load(&p->x); v.y = pub_load(&p->y); return v; }

// Grid barrier number `it` over all G workgroups of a cooperative launch.
// Every wave waits for its outstanding memory operations (the agent-scope
// exchange stores) to complete before the workgroup barrier after which the
// arrival is counted: a workgroup-scope fence alone emits no such wait on
// gfx950, and the stores could otherwise become visible after the arrival.  With more than 64 workgroups, arrivals go
// to 8 group counters (g % 8) and the last arriver of a group bumps the top
// counter, which keeps contention on one address bounded.  All counters
// start at 0.  A wait that exceeds the spin limit, or sees another
// workgroup's failure flag, sets the flag (and the sticky flag, if given)
// and returns, so every workgroup always reaches the end of the kernel.
__device__ inline void grid_barrier(unsigned* bar, unsigned it, int G, int* err, int* sticky = nullptr) {
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned target;
    if (G > 64) {
      const int grp = blockIdx.x & 7;
      const unsigned gsize = (unsigned)((G - grp + 7) / 8);
      const unsigned old = __hip_atomic_fetch_add(bar + 64 * (1 + grp), 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
      if (old + 1 == (it + 1) * gsize) __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      target = (it + 1) * 8u;
    } else {
      __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      target = (it + 1) * (unsigned)G;
    }
    long long spins = 0;
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      if (++spins > kSpinLimit) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (sticky) __hip_atomic_store(sticky, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Leaf LU of an m x w panel (w <= kGetf2MaxW) at A (row-major, lda).  Workgroup g
// owns panel rows [g R, g R + R).  ipiv and info use global indices (row0 is
// the panel's global row, which equals its global column).  The launch
// counts its barriers in wk.bar and clears wk.bar_next for the next leaf.
// If a barrier fails, the panel is written as NaN and info (if still 0)
// becomes MG_ERR_SYNC, so the failure reaches the factors and the caller.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_getf2(typename C2<T>::t* __restrict__ A, int64_t lda, int m, int w,
                                                   int R, int64_t row0, int* __restrict__ ipiv,
                                                   int* __restrict__ info, mg::Getf2Work wk) {
  typedef typename C2<T>::t T2;
  extern __shared__ unsigned char smem_raw[];
  T2* tile = reinterpret_cast<T2*>(smem_raw);
  __shared__ T2 prow[mg::kGetf2MaxW], orow[mg::kGetf2MaxW];
  __shared__ double sv[kThreads / 64];
  __shared__ int sr[kThreads / 64];

  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
  const int rb = g * R, re = min(m, rb + R);
  const int nr = re > rb ? re - rb : 0;
  T2* cdata = reinterpret_cast<T2*>(wk.cand_data);
  T2* rowj = reinterpret_cast<T2*>(wk.rowj);

  int* const lerr = reinterpret_cast<int*>(wk.bar) + kErrWord;
  for (int i = g * kThreads + tid; i < kBarWords; i += G * kThreads) wk.bar_next[i] = 0u;
  for (int e = tid; e < nr * w; e += kThreads) tile[e] = A[(int64_t)(rb + e / w) * lda + e % w];
  __syncthreads();

  for (int j = 0; j < w; ++j) {
    const int par = j & 1;
    double best = -1.0;
    int brow = INT_MAX;
    for (int lr = tid; lr < nr; lr += kThreads) {
      const int grow = rb + lr;
      if (grow >= j) {
        const T2 v = tile[lr * w + j];
        const double a = fabs((double)v.x) + fabs((double)v.y);
        if (a > best) { best = a; brow = grow; }
      }
    }
    block_argmax(best, brow, sv, sr);
    if (tid == 0) {
      pub_store(&wk.cand_val[par * G + g], best);
      pub_store(&wk.cand_row[par * G + g], brow);
    }
    if (brow != INT_MAX && tid < w)
      pub_store2(&cdata[(par * G + g) * mg::kGetf2MaxW + tid], tile[(brow - rb) * w + tid]);
    if (j >= rb && j < re && tid < w) pub_store2(&rowj[par * mg::kGetf2MaxW + tid], tile[(j - rb) * w + tid]);
    if (G > 1) grid_barrier(wk.bar, (unsigned)j, G, lerr, wk.err);
    else __syncthreads();

    double pv = -1.0;
    int p = INT_MAX;
    if (tid < G) { pv = pub_load(&wk.cand_val[par * G + tid]); p = pub_load(&wk.cand_row[par * G + tid]); }
    block_argmax(pv, p, sv, sr);
    if (p == INT_MAX) p = j;   // whole column NaN: keep row j, NaN propagates
    const int pwg = p / R;
    if (tid < w) {
      orow[tid] = pub_load2(&rowj[par * mg::kGetf2MaxW + tid]);
      prow[tid] = (p == j) ? orow[tid] : pub_load2(&cdata[(par * G + pwg) * mg::kGetf2MaxW + tid]);
    }
    __syncthreads();
    if (tid < w) {
      if (j >= rb && j < re) tile[(j - rb) * w + tid] = prow[tid];
      if (p != j && p >= rb && p < re) tile[(p - rb) * w + tid] = orow[tid];
    }
    if (g == 0 && tid == 0) {
      ipiv[row0 + j] = (int)(row0 + p);
      if (pv == 0.0) atomicCAS(info, 0, (int)(row0 + j + 1));
    }
    __syncthreads();
    const T2 piv = prow[j];
    if (piv.x != 0 || piv.y != 0) {
      const T2 inv = crecip(piv);
      for (int lr = tid; lr < nr; lr += kThreads)
        if (rb + lr > j) tile[lr * w + j] = cmul(tile[lr * w + j], inv);
    }
    __syncthreads();
    const int wr = w - j - 1;
    if (wr > 0) {
      for (int e = tid; e < nr * wr; e += kThreads) {
        const int lr = e / wr, c = j + 1 + e % wr;
        if (rb + lr > j) {
          const T2 l = tile[lr * w + j];
          T2 x = tile[lr * w + c];
          const T2 u = prow[c];
          x.x -= l.x * u.x - l.y * u.y;
          x.y -= l.x * u.y + l.y * u.x;
          tile[lr * w + c] = x;
        }
      }
    }
    __syncthreads();
  }
  if (G > 1 && __hip_atomic_load(lerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
    const T nan = __builtin_nan("");
    for (int e = tid; e < nr * w; e += kThreads) tile[e] = T2{nan, nan};
    if (g == 0 && tid == 0) atomicCAS(info, 0, MG_ERR_SYNC);
    __syncthreads();
  }
  for (int e = tid; e < nr * w; e += kThreads) A[(int64_t)(rb + e / w) * lda + e % w] = tile[e];
}

__global__ void k_zero_ints(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Apply the row swaps k1..k2-1 (in order) to columns [c0, c1).
template <typename T2>
__global__ __launch_bounds__(256) void k_laswp(T2* __restrict__ A, int64_t lda, int64_t c0, int64_t c1, int64_t k1,
                                                int64_t k2, const int* __restrict__ ipiv) {
  const int64_t c = c0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= c1) return;
  for (int64_t i = k1; i < k2; ++i) {
    const int64_t p = ipiv[i];
    if (p != i) {
      const T2 a = A[i * lda + c];
      A[i * lda + c] = A[p * lda + c];
      A[p * lda + c] = a;
    }
  }
}

// Triangular solve of an nb x nb (nb <= 32) block against ncols right-hand
// sides: unit-lower forward or non-unit upper backward substitution.  Each
// half-wave holds one column with one row per lane; the solved x_i is
// broadcast by a lane shuffle and the other rows update in parallel, so a
// column takes nb dependent steps of one complex FMA each.  A workgroup of
// 256 threads covers 8 columns.
template <typename T, bool UPPER>
__global__ __launch_bounds__(256) void k_trsm_leaf(const typename C2<T>::t* __restrict__ L, int64_t ldl, int64_t sL,
                                                   int nb, typename C2<T>::t* __restrict__ B, int64_t ldb,
                                                   int64_t sB, int64_t ncols) {
  typedef typename C2<T>::t T2;
  L += blockIdx.y * sL;   // batch
  B += blockIdx.y * sB;
  __shared__ T2 Ls[mg::kLeaf][mg::kLeaf + 1];
  __shared__ T2 dinv[mg::kLeaf];
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) Ls[e / nb][e % nb] = L[(e / nb) * ldl + e % nb];
  __syncthreads();
  if (UPPER && threadIdx.x < nb) dinv[threadIdx.x] = crecip(Ls[threadIdx.x][threadIdx.x]);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int r = lane & 31, half = lane & 32;
  const int64_t c = (int64_t)blockIdx.x * 8 + (threadIdx.x >> 5);
  const bool live = c < ncols && r < nb;
  T2 x;
  x.x = 0; x.y = 0;
  if (live) x = B[r * ldb + c];
  if (!UPPER) {
    for (int i = 0; i < nb - 1; ++i) {
      T2 xi;
      xi.x = __shfl(x.x, half + i);
      xi.y = __shfl(x.y, half + i);
      if (r > i && r < nb) {
        const T2 l = Ls[r][i];
        x.x -= l.x * xi.x - l.y * xi.y;
        x.y -= l.x * xi.y + l.y * xi.x;
      }
    }
  } else {
    for (int i = nb - 1; i >= 0; --i) {
      if (r == i) x = cmul(x, dinv[i]);
      T2 xi;
      xi.x = __shfl(x.x, half + i);
      xi.y = __shfl(x.y, half + i);
      if (r < i) {
        const T2 u = Ls[r][i];
        x.x -= u.x * xi.x - u.y * xi.y;
        x.y -= u.x * xi.y + u.y * xi.x;
      }
    }
  }
  if (live) B[r * ldb + c] = x;
}

inline int64_t split_half(int64_t w) {
  // left part: half, rounded up to a multiple of 16 (keeps GEMM tiles whole)
  int64_t h = ((w / 2) + 15) / 16 * 16;
  if (h >= w) h = w / 2;
  return h > 0 ? h : 1;
}

}  // namespace

namespace mg {

template <typename T>
int laswp(void* A, int64_t lda, int64_t c0, int64_t c1, int64_t k1, int64_t k2, const int* ipiv, hipStream_t s) {
  typedef typename C2<T>::t T2;
  if (c1 <= c0 || k2 <= k1) return MG_OK;
  const int64_t nb = (c1 - c0 + 255) / 256;
  hipLaunchKernelGGL(k_laswp<T2>, dim3((unsigned)nb), dim3(256), 0, s, reinterpret_cast<T2*>(A), lda, c0, c1, k1, k2,
                     ipiv);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

// Recursive triangular solves, batched: problem b uses L + b sL and B + b sB.
template <typename T>
int trsm_lower_unit_b(const void* L, int64_t ldl, int64_t sL, int64_t nb, void* B, int64_t ldb, int64_t sB,
                             int64_t ncols, int64_t batch, hipStream_t s) {
  typedef typename C2<T>::t T2;
  if (nb <= 0 || ncols <= 0 || batch <= 0) return MG_OK;
  if (nb <= kLeaf) {
    hipLaunchKernelGGL((k_trsm_leaf<T, false>), dim3((unsigned)((ncols + 7) / 8), (unsigned)batch), dim3(256), 0, s,
                       reinterpret_cast<const T2*>(L), ldl, sL, (int)nb, reinterpret_cast<T2*>(B), ldb, sB, ncols);
    MG_HIP(hipGetLastError());
    return MG_OK;
  }
  const int64_t n1 = split_half(nb);
  const T2* Lp = reinterpret_cast<const T2*>(L);
  T2* Bp = reinterpret_cast<T2*>(B);
  MG_TRY(trsm_lower_unit_b<T>(Lp, ldl, sL, n1, Bp, ldb, sB, ncols, batch, s));
  MG_TRY(cgemm_batched<T>(nb - n1, ncols, n1, cd{-1, 0}, Lp + n1 * ldl, ldl, sL, Bp, ldb, sB, cd{1, 0},
                          Bp + n1 * ldb, ldb, sB, batch, s));
  return trsm_lower_unit_b<T>(Lp + n1 * ldl + n1, ldl, sL, nb - n1, Bp + n1 * ldb, ldb, sB, ncols, batch, s);
}

template <typename T>
int trsm_upper_b(const void* U, int64_t ldu, int64_t sU, int64_t nb, void* B, int64_t ldb, int64_t sB,
                        int64_t ncols, int64_t batch, hipStream_t s) {
  typedef typename C2<T>::t T2;
  if (nb <= 0 || ncols <= 0 || batch <= 0) return MG_OK;
  if (nb <= kLeaf) {
    hipLaunchKernelGGL((k_trsm_leaf<T, true>), dim3((unsigned)((ncols + 7) / 8), (unsigned)batch), dim3(256), 0, s,
                       reinterpret_cast<const T2*>(U), ldu, sU, (int)nb, reinterpret_cast<T2*>(B), ldb, sB, ncols);
    MG_HIP(hipGetLastError());
    return MG_OK;
  }
  const int64_t n1 = split_half(nb);
  const T2* Up = reinterpret_cast<const T2*>(U);
  T2* Bp = reinterpret_cast<T2*>(B);
  MG_TRY(trsm_upper_b<T>(Up + n1 * ldu + n1, ldu, sU, nb - n1, Bp + n1 * ldb, ldb, sB, ncols, batch, s));
  MG_TRY(cgemm_batched<T>(n1, ncols, nb - n1, cd{-1, 0}, Up + n1, ldu, sU, Bp + n1 * ldb, ldb, sB, cd{1, 0}, Bp,
                          ldb, sB, batch, s));
  return trsm_upper_b<T>(Up, ldu, sU, n1, Bp, ldb, sB, ncols, batch, s);
}

template <typename T>
int trsm_lower_unit(const void* L, int64_t ldl, int64_t nb, void* B, int64_t ldb, int64_t ncols, hipStream_t s) {
  return trsm_lower_unit_b<T>(L, ldl, 0, nb, B, ldb, 0, ncols, 1, s);
}

template <typename T>
int trsm_upper(const void* U, int64_t ldu, int64_t nb, void* B, int64_t ldb, int64_t ncols, hipStream_t s) {
  return trsm_upper_b<T>(U, ldu, 0, nb, B, ldb, 0, ncols, 1, s);
}

// Identity into `count` nb x nb blocks (ld nb, consecutive).
template <typename T2>
__global__ void k_eye_blocks(T2* D, int nb, int64_t count) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count * nb * nb) return;
  const int64_t w = e % ((int64_t)nb * nb);
  T2 v;
  v.x = (w / nb == w % nb) ? 1 : 0;
  v.y = 0;
  D[e] = v;
}

// Inverses of the kInvBlock x kInvBlock diagonal blocks of L (unit lower) and
// U: dinv[0 .. nblk) hold inv(L_bb), dinv[nblk .. 2 nblk) inv(U_bb), each a
// row-major kInvBlock^2 block (the last, partial block uses its top-left).
template <typename T>
int diag_inverses(const void* LU, int64_t n, int64_t lda, void* dinv, hipStream_t s) {
  typedef typename C2<T>::t T2;
  if (n <= 0) return MG_OK;
  const int64_t nb = kInvBlock, nblk = (n + nb - 1) / nb, full = n / nb, rem = n - full * nb;
  const int64_t bs = nb * nb;
  T2* D = reinterpret_cast<T2*>(dinv);
  const T2* A = reinterpret_cast<const T2*>(LU);
  const int64_t tot = 2 * nblk * bs;
  hipLaunchKernelGGL(k_eye_blocks<T2>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, D, (int)nb, 2 * nblk);
  MG_HIP(hipGetLastError());
  const int64_t sdiag = nb * lda + nb;   // stride between diagonal blocks of LU
  MG_TRY(trsm_lower_unit_b<T>(A, lda, sdiag, nb, D, nb, bs, nb, full, s));
  MG_TRY(trsm_upper_b<T>(A, lda, sdiag, nb, D + nblk * bs, nb, bs, nb, full, s));
  if (rem) {
    const T2* Ab = A + full * sdiag;
    MG_TRY(trsm_lower_unit_b<T>(Ab, lda, 0, rem, D + full * bs, nb, 0, rem, 1, s));
    MG_TRY(trsm_upper_b<T>(Ab, lda, 0, rem, D + (nblk + full) * bs, nb, 0, rem, 1, s));
  }
  return MG_OK;
}

// Solve A X = B (B n x nrhs, overwritten) from getrf's factors and
// diag_inverses' blocks, using only GEMMs: forward W_b = inv(L_bb) B_b,
// B_{>b} -= L_{>b,b} W_b; backward B_b = inv(U_bb) W_b, W_{<b} -= U_{<b,b} B_b.
// work: n x nrhs scratch (ld nrhs).
template <typename T>
int getrs_dinv(const void* LU, int64_t n, int64_t lda, const int* ipiv, const void* dinv, void* B, int64_t ldb,
               int64_t nrhs, void* work, hipStream_t s, GemmFn gemm) {
  typedef typename C2<T>::t T2;
  if (n < 0 || nrhs < 0 || lda < n || ldb < nrhs) return MG_ERR_ARG;
  if (n == 0 || nrhs == 0) return MG_OK;
  const GemmFn g = gemm ? gemm : &cgemm<T>;
  const int64_t nb = kInvBlock, nblk = (n + nb - 1) / nb, bs = nb * nb;
  const T2* A = reinterpret_cast<const T2*>(LU);
  const T2* D = reinterpret_cast<const T2*>(dinv);
  T2* Bp = reinterpret_cast<T2*>(B);
  T2* W = reinterpret_cast<T2*>(work);
  const int64_t ldw = nrhs;
  if (ipiv) MG_TRY(laswp<T>(B, ldb, 0, nrhs, 0, n, ipiv, s));   // no ipiv: unpivoted factors
  for (int64_t b = 0; b < nblk; ++b) {
    const int64_t r0 = b * nb, r1 = (r0 + nb < n) ? r0 + nb : n, h = r1 - r0;
    MG_TRY(g(h, nrhs, h, cd{1, 0}, D + b * bs, nb, Bp + r0 * ldb, ldb, cd{0, 0}, W + r0 * ldw, ldw, s));
    if (r1 < n)
      MG_TRY(g(n - r1, nrhs, h, cd{-1, 0}, A + r1 * lda + r0, lda, W + r0 * ldw, ldw, cd{1, 0}, Bp + r1 * ldb, ldb,
               s));
  }
  for (int64_t b = nblk - 1; b >= 0; --b) {
    const int64_t r0 = b * nb, r1 = (r0 + nb < n) ? r0 + nb : n, h = r1 - r0;
    MG_TRY(g(h, nrhs, h, cd{1, 0}, D + (nblk + b) * bs, nb, W + r0 * ldw, ldw, cd{0, 0}, Bp + r0 * ldb, ldb, s));
    if (r0 > 0) MG_TRY(g(r0, nrhs, h, cd{-1, 0}, A + r0, lda, Bp + r0 * ldb, ldb, cd{1, 0}, W, ldw, s));
  }
  return MG_OK;
}

// rows per leaf workgroup to aim for (MG_GETF2_ROWS overrides kGetf2MinRows)
static int getf2_min_rows() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("MG_GETF2_ROWS");
    const int r = e ? atoi(e) : 0;
    v = r >= 16 && r <= 4096 ? r : kGetf2MinRows;
  }
  return v;
}

template <typename T>
static int getf2(Workspace& ws, void* A, int64_t lda, int64_t k, int64_t m, int64_t w, int* ipiv, int* info,
                 hipStream_t s) {
  typedef typename C2<T>::t T2;
  const int64_t bytes_per_row = w * (int64_t)sizeof(T2);
  const int64_t max_rows = kGetf2Lds / bytes_per_row;
  const int min_rows = getf2_min_rows();
  int G = (int)((m + min_rows - 1) / min_rows);
  const int gmax = ws.coop_blocks;
  if (G > gmax) G = gmax;
  if (G < 1) G = 1;
  int R = (int)((m + G - 1) / G);
  if (R > max_rows) {
    G = (int)((m + max_rows - 1) / max_rows);
    if (G > gmax) {
      set_error("getf2: panel of %lld rows does not fit %d workgroups", (long long)m, gmax);
      return MG_ERR_UNSUPPORTED;
    }
    R = (int)((m + G - 1) / G);
  }
  T2* Ap = reinterpret_cast<T2*>(A) + k * lda + k;
  int mi = (int)m, wi = (int)w;
  // alternate barrier sets: this launch counts in one and clears the other
  // (leaves of a handle are stream-ordered, so the previous leaf is done)
  Getf2Work wk = ws.getf2;
  wk.bar = ws.getf2.bar + ws.leaf_parity * kBarWords;
  wk.bar_next = ws.getf2.bar + (1 - ws.leaf_parity) * kBarWords;
  const size_t shmem = (size_t)R * bytes_per_row;
  // G <= the CU count, one workgroup per CU: every workgroup becomes resident
  // once concurrently running kernels (which never wait on this one) drain,
  // and each barrier wait is bounded by the spin limit.
  hipLaunchKernelGGL(k_getf2<T>, dim3(G), dim3(kThreads), shmem, s, Ap, lda, mi, wi, R, k, ipiv, info, wk);
  MG_HIP(hipGetLastError());
  ws.leaf_parity ^= 1;       // only a launched leaf has cleared the other set
  return MG_OK;
}

template <typename T>
static int panel(Workspace& ws, void* A, int64_t lda, int64_t k, int64_t m, int64_t w, int* ipiv, int* info,
                 hipStream_t s) {
  typedef typename C2<T>::t T2;
  // leaf width: the widest power of two up to kGetf2MaxW whose rows fit the
  // LDS of the workgroups a leaf may use (the narrower the leaf, the more
  // launches of swaps, triangular solves and GEMMs the recursion adds)
  int64_t lw = kGetf2MaxW;
  while (lw > 1 && (int64_t)ws.coop_blocks * (kGetf2Lds / (lw * (int64_t)sizeof(T2))) < m) lw /= 2;
  if (w <= lw) return getf2<T>(ws, A, lda, k, m, w, ipiv, info, s);
  T2* Ap = reinterpret_cast<T2*>(A);
  const int64_t w1 = split_half(w);
  MG_TRY(panel<T>(ws, A, lda, k, m, w1, ipiv, info, s));
  MG_TRY(laswp<T>(A, lda, k + w1, k + w, k, k + w1, ipiv, s));
  MG_TRY(trsm_lower_unit<T>(Ap + k * lda + k, lda, w1, Ap + k * lda + k + w1, lda, w - w1, s));
  MG_TRY(cgemm<T>(m - w1, w - w1, w1, cd{-1, 0}, Ap + (k + w1) * lda + k, lda, Ap + k * lda + k + w1, lda, cd{1, 0},
                  Ap + (k + w1) * lda + k + w1, lda, s));
  MG_TRY(panel<T>(ws, A, lda, k + w1, m - w1, w - w1, ipiv, info, s));
  return laswp<T>(A, lda, k, k + w1, k + w1, k + w, ipiv, s);
}

template <typename T>
int getrf(Workspace& ws, void* A, int64_t n, int64_t lda, int* ipiv, int* info, hipStream_t s, int nb,
          GemmFn update) {
  typedef typename C2<T>::t T2;
  if (n < 0 || lda < n) return MG_ERR_ARG;
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, info, 1);
  if (n == 0) return MG_OK;
  T2* Ap = reinterpret_cast<T2*>(A);
  GemmFn fn = update ? update : &cgemm_trail<T>;
  // C[r0:, c0:c0+nc] -= L[r0:, k:k+jb] U[k:k+jb, c0:c0+nc], timed when profiling
  auto trail = [&](int64_t k, int64_t jb, int64_t c0, int64_t nc) -> int {
    const int64_t r0 = k + jb, r = n - r0;
    if (r <= 0 || nc <= 0) return MG_OK;
    return timed_update(ws, fn, r, nc, jb, cd{-1, 0}, Ap + r0 * lda + k, lda, Ap + k * lda + c0, lda, cd{1, 0},
                        Ap + r0 * lda + c0, lda, s);
  };
  // Look-ahead of depth one: once the next panel's columns are updated, that
  // panel is factored on the workspace's second stream while the rest of the
  // trailing update runs here.  The two touch disjoint columns; the next
  // step waits for the panel before applying its row swaps.
  MG_TRY(panel<T>(ws, A, lda, 0, n, (n < nb) ? n : nb, ipiv, info, s));
  for (int64_t k = 0; k < n; k += nb) {
    const int64_t jb = (n - k < nb) ? n - k : nb;
    const int64_t kn = k + jb;
    MG_TRY(laswp<T>(A, lda, kn, n, k, kn, ipiv, s));
    if (kn < n) {
      const int64_t jb2 = (n - kn < nb) ? n - kn : nb;
      MG_TRY(trsm_lower_unit<T>(Ap + k * lda + k, lda, jb, Ap + k * lda + kn, lda, n - kn, s));
      MG_TRY(trail(k, jb, kn, jb2));
      if (ws.s2) {
        MG_HIP(hipEventRecord(ws.evA, s));
        MG_HIP(hipStreamWaitEvent(ws.s2, ws.evA, 0));
        MG_TRY(panel<T>(ws, A, lda, kn, n - kn, jb2, ipiv, info, ws.s2));
        MG_HIP(hipEventRecord(ws.evB, ws.s2));
        MG_TRY(trail(k, jb, kn + jb2, n - kn - jb2));
        MG_HIP(hipStreamWaitEvent(s, ws.evB, 0));
      } else {
        MG_TRY(trail(k, jb, kn + jb2, n - kn - jb2));
        MG_TRY(panel<T>(ws, A, lda, kn, n - kn, jb2, ipiv, info, s));
      }
    }
    MG_TRY(laswp<T>(A, lda, 0, k, k, kn, ipiv, s));
  }
  return MG_OK;
}

template <typename T>
int getrs(const void* LU, int64_t n, int64_t lda, const int* ipiv, void* B, int64_t ldb, int64_t nrhs,
          hipStream_t s) {
  if (n < 0 || nrhs < 0 || lda < n || ldb < nrhs) return MG_ERR_ARG;
  if (n == 0 || nrhs == 0) return MG_OK;
  MG_TRY(laswp<T>(B, ldb, 0, nrhs, 0, n, ipiv, s));
  MG_TRY(trsm_lower_unit<T>(LU, lda, n, B, ldb, nrhs, s));
  return trsm_upper<T>(LU, lda, n, B, ldb, nrhs, s);
}

int workspace_init(Workspace& ws) {
  if (ws.ready) return MG_OK;
  int dev = 0;
  MG_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  MG_HIP(hipGetDeviceProperties(&prop, dev));
  int per_cu_d = 0, per_cu_f = 0;
  MG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_d, reinterpret_cast<const void*>(&k_getf2<double>),
                                                      kThreads, kGetf2Lds));
  MG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_f, reinterpret_cast<const void*>(&k_getf2<float>),
                                                      kThreads, kGetf2Lds));
  int per_cu = per_cu_d < per_cu_f ? per_cu_d : per_cu_f;
  if (per_cu < 1) per_cu = 1;
  // one workgroup per CU at most: the barrier cost grows with the count
  ws.coop_blocks = prop.multiProcessorCount;
  ws.cus = prop.multiProcessorCount;
  (void)per_cu;
  const int G = kMaxGetf2Blocks;
  MG_HIP(hipMalloc(&ws.getf2.cand_val, sizeof(double) * 2 * G));
  MG_HIP(hipMalloc(&ws.getf2.cand_row, sizeof(int) * 2 * G));
  MG_HIP(hipMalloc(&ws.getf2.cand_data, sizeof(double2) * 2 * G * kGetf2MaxW));
  MG_HIP(hipMalloc(&ws.getf2.rowj, sizeof(double2) * 2 * kGetf2MaxW));
  MG_HIP(hipMalloc(&ws.getf2.bar, sizeof(unsigned) * 2 * kBarWords));
  MG_HIP(hipMemset(ws.getf2.bar, 0, sizeof(unsigned) * 2 * kBarWords));
  MG_HIP(hipMalloc(&ws.inv_exch, sizeof(float2) * 2 * 256 * 32));
  MG_HIP(hipMalloc(&ws.inv_ctl, sizeof(int) * 128));
  MG_HIP(hipMemset(ws.inv_ctl, 0, sizeof(int) * 128));
  MG_HIP(hipMalloc(&ws.getf2.err, sizeof(int) * 64));
  MG_HIP(hipMemset(ws.getf2.err, 0, sizeof(int) * 64));
  if (ws.coop_blocks > G) ws.coop_blocks = G;
  MG_HIP(hipStreamCreateWithFlags(&ws.s2, hipStreamNonBlocking));
  MG_HIP(hipStreamCreateWithFlags(&ws.s3, hipStreamNonBlocking));
  MG_HIP(hipStreamCreateWithFlags(&ws.s4, hipStreamNonBlocking));
  MG_HIP(hipEventCreateWithFlags(&ws.evA, hipEventDisableTiming));
  MG_HIP(hipEventCreateWithFlags(&ws.evB, hipEventDisableTiming));
  MG_HIP(hipEventCreateWithFlags(&ws.evC, hipEventDisableTiming));
  MG_HIP(hipEventCreateWithFlags(&ws.evD, hipEventDisableTiming));
  MG_HIP(hipEventCreateWithFlags(&ws.evE, hipEventDisableTiming));
  MG_HIP(hipEventCreateWithFlags(&ws.evF, hipEventDisableTiming));
  MG_HIP(hipEventCreateWithFlags(&ws.evG, hipEventDisableTiming));
  MG_HIP(hipEventCreateWithFlags(&ws.evH, hipEventDisableTiming));
  ws.ready = true;
  return MG_OK;
}

int profile_read(Workspace& ws, hipStream_t s, long long* launches, double* flops, double* ms, bool per_class) {
  MG_HIP(hipStreamSynchronize(s));
  double f[2] = {0.0, 0.0}, t[2] = {0.0, 0.0};
  long long c[2] = {0, 0};
  for (size_t i = 0; i < ws.prof.used; ++i) {
    float e = 0.f;
    MG_HIP(hipEventElapsedTime(&e, ws.prof.ev[2 * i], ws.prof.ev[2 * i + 1]));
    const int k = ws.prof.cls[i] ? 1 : 0;
    t[k] += e;
    f[k] += ws.prof.flops[i];
    ++c[k];
  }
  for (int k = 0; k < (per_class ? 2 : 1); ++k) {
    const int q = per_class ? k : 0;
    if (launches) launches[q] = per_class ? c[k] : c[0] + c[1];
    if (flops) flops[q] = per_class ? f[k] : f[0] + f[1];
    if (ms) ms[q] = per_class ? t[k] : t[0] + t[1];
  }
  ws.prof.used = 0;
  return MG_OK;
}

void workspace_free(Workspace& ws) {
  for (hipEvent_t e : ws.prof.ev) (void)hipEventDestroy(e);
  ws.prof.ev.clear();
  if (!ws.ready) return;
  (void)hipFree(ws.getf2.cand_val);
  (void)hipFree(ws.getf2.cand_row);
  (void)hipFree(ws.getf2.cand_data);
  (void)hipFree(ws.getf2.rowj);
  (void)hipFree(ws.getf2.bar);
  (void)hipFree(ws.getf2.err);
  if (ws.sweep) (void)hipFree(ws.sweep);
  if (ws.inv_exch) (void)hipFree(ws.inv_exch);
  if (ws.inv_ctl) (void)hipFree(ws.inv_ctl);
  if (ws.s2) (void)hipStreamDestroy(ws.s2);
  if (ws.s3) (void)hipStreamDestroy(ws.s3);
  if (ws.s4) (void)hipStreamDestroy(ws.s4);
  if (ws.evA) (void)hipEventDestroy(ws.evA);
  if (ws.evB) (void)hipEventDestroy(ws.evB);
  if (ws.evC) (void)hipEventDestroy(ws.evC);
  if (ws.evD) (void)hipEventDestroy(ws.evD);
  if (ws.evE) (void)hipEventDestroy(ws.evE);
  if (ws.evF) (void)hipEventDestroy(ws.evF);
  if (ws.evG) (void)hipEventDestroy(ws.evG);
  if (ws.evH) (void)hipEventDestroy(ws.evH);
  ws = Workspace();
}

int timed_update(Workspace& ws, GemmFn fn, int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda,
                 const void* B, int64_t ldb, cd beta, void* C, int64_t ldc, hipStream_t s) {
  const int cls = (fn == &cgemm_trail<double> && zgemm_uses_dma(M, N, K)) ? 1 : 0;
  return timed_launch(
      ws, 8.0 * (double)M * (double)N * (double)K, s, [&] { return fn(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s); },
      cls);
}

int workspace_leaf_blocks(Workspace& ws, int g) {
  const int top = (ws.cus < kMaxGetf2Blocks) ? ws.cus : kMaxGetf2Blocks;
  ws.coop_blocks = (g <= 0 || g > top) ? top : g;
  return MG_OK;
}

int workspace_check(Workspace& ws, hipStream_t s) {
  if (!ws.ready) return MG_OK;
  int err = 0;
  MG_HIP(hipMemcpyAsync(&err, ws.getf2.err, sizeof(int), hipMemcpyDeviceToHost, s));
  MG_HIP(hipStreamSynchronize(s));
  if (err) {
    // reported once: the next factorisation on this handle starts clean
    // (each leaf launch has its own failure word)
    MG_HIP(hipMemsetAsync(ws.getf2.err, 0, sizeof(int), s));
    MG_HIP(hipStreamSynchronize(s));
    set_error("a grid barrier of the LU leaf did not complete");
    return MG_ERR_SYNC;
  }
  return MG_OK;
}

template int laswp<double>(void*, int64_t, int64_t, int64_t, int64_t, int64_t, const int*, hipStream_t);
template int laswp<float>(void*, int64_t, int64_t, int64_t, int64_t, int64_t, const int*, hipStream_t);
template int trsm_lower_unit<double>(const void*, int64_t, int64_t, void*, int64_t, int64_t, hipStream_t);
template int trsm_lower_unit<float>(const void*, int64_t, int64_t, void*, int64_t, int64_t, hipStream_t);
template int trsm_upper<double>(const void*, int64_t, int64_t, void*, int64_t, int64_t, hipStream_t);
template int trsm_upper<float>(const void*, int64_t, int64_t, void*, int64_t, int64_t, hipStream_t);
template int getrf<double>(Workspace&, void*, int64_t, int64_t, int*, int*, hipStream_t, int, GemmFn);
template int getrf<float>(Workspace&, void*, int64_t, int64_t, int*, int*, hipStream_t, int, GemmFn);
template int getrs<double>(const void*, int64_t, int64_t, const int*, void*, int64_t, int64_t, hipStream_t);
template int getrs<float>(const void*, int64_t, int64_t, const int*, void*, int64_t, int64_t, hipStream_t);
template int trsm_lower_unit_b<double>(const void*, int64_t, int64_t, int64_t, void*, int64_t, int64_t, int64_t, int64_t,
                                       hipStream_t);
template int trsm_lower_unit_b<float>(const void*, int64_t, int64_t, int64_t, void*, int64_t, int64_t, int64_t, int64_t,
                                      hipStream_t);
template int diag_inverses<double>(const void*, int64_t, int64_t, void*, hipStream_t);
template int diag_inverses<float>(const void*, int64_t, int64_t, void*, hipStream_t);
template int getrs_dinv<double>(const void*, int64_t, int64_t, const int*, const void*, void*, int64_t, int64_t, void*,
                                hipStream_t, GemmFn);
template int getrs_dinv<float>(const void*, int64_t, int64_t, const int*, const void*, void*, int64_t, int64_t, void*,
                               hipStream_t, GemmFn);

}  // namespace mg
