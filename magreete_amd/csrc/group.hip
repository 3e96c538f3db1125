// Grouped LU factorisation and solve of many independent complex matrices of
// uneven order, in lockstep: every launch of the blocked algorithm serves
// every problem of the group that still has work at that column.
//
// An ensemble of small systems solved one at a time (magreete.py:346-452,
// one torch.linalg.solve per configuration, Transmission3D.py:205) is bound
// by the dependency chain of each factorisation: the pivot search of every
// column, the small triangular solves and the small products inside each
// panel.  Here the problems of a group advance column by column together, so
// the chain is paid once per group instead of once per problem, and every
// trailing update, panel product and triangular solve is one launch over a
// grouped tile table.
//
// Layout: the problems are sorted by decreasing order; problem p owns its
// own row-major matrix (A, lda, n), pivots ipiv (n ints, 0-based, LAPACK's
// meaning), info (first zero pivot, 1-based) and right-hand sides (B, ldb,
// nrhs columns).  The table of these descriptors lives on the device.  A
// launch serves the problems p < np (those with work at its column); problem
// p owns workgroups [first[p], first[p+1]) of the launch, and the tile table
// (first[]) travels in the kernel arguments, so a captured graph carries it.
//
// Algorithm (per problem the same as getrf in lu.hip, with its pivoting
// rule: the largest |re| + |im| on or below the diagonal, lowest row on
// ties): right-looking with panels of nb columns; a panel is factored
// recursively (left half, row swaps + triangular solve + product on the
// right half, right half) down to leaves of at most kGLeaf columns.  A leaf
// is factored by one workgroup per problem that streams the leaf's rows
// through the caches, one pass per column (rank-1 update by the previous
// pivot, scaling of the previous column and this column's pivot search
// fused), so no workgroup ever waits for another and any number of problems
// can share a launch.
#include "cgemm.h"
#include "lu.h"
#include "mg_internal.h"
#include "dma_update.h"
#include "psplit.h"
#include "split_gemm.h"
#include <algorithm>
#include <climits>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

namespace mg {

constexpr int kMaxGroup = 256;   // problems of one group

struct GrpProb {
  void* A;        // the matrix (factored in place)
  int64_t lda;
  int64_t n;
  int* ipiv;
  int* info;
  void* B;        // right-hand sides (n x nrhs)
  int64_t ldb;
  void* X;        // auxiliary n x nrhs block: the solution of the residual product B -= A X
  int64_t ldx;
  int64_t idx;    // the caller's index of this problem
  void* SA[2];    // BF16x3: the pre-split A operand (L21) of even / odd steps
  void* SB;       // BF16x3: the pre-split B operand (U12) of the current step
  int* perm;      // BF16x3 solves: position r of P A holds row perm[r] of A
  void* Dinv;     // BF16x3 solves: the inverted 64 x 64 diagonal blocks of L and U
  void* W1;       // BF16x3 solves: n x nrhs, the permuted right-hand sides (forward sweep in place)
  void* W2;       // BF16x3 solves: n x nrhs, the forward sweep's result (backward sweep in place)
};

// Workgroups of one grouped launch: problem p < np owns [first[p], first[p+1]).
struct GrpMap {
  int np;
  int first[kMaxGroup + 1];
};

// Tunables of a group's factorisation.
struct GrpCfg {
  int nb = 128;      // panel width (trailing-update depth)
  int leaf = 16;     // leaf width, <= kGLeaf
  int tile = 0;      // trailing-update tile: 0 = 128 x 64 (three-product), 1 = 64 x 64
  int lthreads = 1024; // threads of a leaf workgroup (the leaf streams its rows: more lanes, more loads in flight)
  int pivot = 1;       // BF16x3 leaves: 1 = tournament pivoting, 0 = partial pivoting (LAPACK's pivots)
};

}  // namespace mg

namespace {

using mg::GrpMap;
using mg::GrpProb;

constexpr int kGLeaf = 16;                    // widest leaf (16 lanes per row)
constexpr int kGUnroll = 4;                   // rows in flight per lane
constexpr int kGTri = 64;                     // widest triangular block of the grouped solves

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
__device__ inline T2 zero2() {
  T2 z;
  z.x = 0;
  z.y = 0;
  return z;
}

template <typename T2>
__device__ inline T2 cmul(T2 a, T2 b) {
  T2 r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

__device__ inline bool better(double v, int r, double bv, int br) { return v > bv || (v == bv && r < br); }

// The problem owning workgroup `bid`: the last p with first[p] <= bid (the
// first[] of problems without workgroups repeat their successor's).
__device__ inline int grp_find(const GrpMap& m, int bid) {
  int lo = 0, hi = m.np;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (m.first[mid] <= bid) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Workgroup ids are dealt round-robin over the 8 XCDs: renumber the leading
// multiple of 8 so that consecutive ids (one problem's tiles) share an XCD.
__device__ inline int xcd_local(int bid, int total) {
  const int per = total / 8;
  return (total >= 64 && bid < 8 * per) ? (bid % 8) * per + bid / 8 : bid;
}

__device__ inline int64_t clamp_extent(int64_t end, int64_t n, int64_t start, bool clamp) {
  const int64_t e = clamp ? (end < n ? end : n) : end;
  return e > start ? e - start : 0;
}

// ---- leaf: one workgroup per problem ---------------------------------------
// Leaf LU of rows [k, n) x columns [k, k + w') of problem blockIdx.x, w' =
// min(w, n - k), by one workgroup of NT threads.  Lane c of each 16-lane
// group holds column c of one row.  Pass j (0 <= j <= w'): rows r >= j get
// column j-1 scaled by the inverse of pivot j-1 and columns >= j updated by
// it, and column j's pivot search runs on the updated values; then rows j
// and the pivot row swap within the leaf.
// The first LR rows of the leaf stay in LDS and the next RS * QR in
// registers (slot q of the lane group rr: row LR + rr + RS q) for all
// passes; only the rows beyond stream through the caches once per pass.
// (144 KiB of LDS; 32 VGPRs of slots at 1024 threads, where a lane has 128
// VGPRs, 40 below)
template <typename T, int NT> struct LeafCache;
template <int NT> struct LeafCache<double, NT> { static constexpr int LR = 576, QR = NT == 1024 ? 8 : 10; };
template <int NT> struct LeafCache<float, NT> { static constexpr int LR = 1152, QR = NT == 1024 ? 16 : 20; };

// One row of a leaf pass, held by a 16-lane group (lane c: column c): the
// update by pivot j-1 (multiplier from lane j-1, scaled by the pivot's
// inverse) and the search of column j.
template <typename T2>
__device__ inline void leaf_row(T2& x, int r, bool valid, bool upd, bool nz, T2 inv, T2 u, int c, int j, int gbase,
                                double& best, int& brow) {
  if (upd) {
    T2 l;
    l.x = __shfl(x.x, gbase + j - 1);
    l.y = __shfl(x.y, gbase + j - 1);
    if (nz) l = cmul(l, inv);
    if (c == j - 1) {
      x = l;
    } else if (c >= j) {
      x.x -= l.x * u.x - l.y * u.y;
      x.y -= l.x * u.y + l.y * u.x;
    }
  }
  if (c == j && valid) {
    const double a = fabs((double)x.x) + fabs((double)x.y);
    if (a > best) {
      best = a;
      brow = r;
    }
  }
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_grp_getf2(const GrpProb* __restrict__ tab, int64_t k, int w) {
  typedef typename C2<T>::t T2;
  constexpr int RS = NT / kGLeaf;   // rows per sweep
  constexpr int LR = LeafCache<T, NT>::LR, QR = LeafCache<T, NT>::QR, CR = LR + RS * QR;
  __shared__ T2 cache[LR][kGLeaf];
  __shared__ T2 urow[kGLeaf], rowa[kGLeaf], rowb[kGLeaf];
  __shared__ T2 sinv;
  __shared__ int snz;
  __shared__ double sv[NT / 64];
  __shared__ int sr[NT / 64];
  const GrpProb pb = tab[blockIdx.x];
  if (pb.n <= k) return;
  const int m = (int)(pb.n - k);
  const int wp = (int)(w < pb.n - k ? w : pb.n - k);
  const int mc = m < LR ? m : LR, mr = m < CR ? m : CR;
  const int64_t lda = pb.lda;
  gptr<T2> A = (gptr<T2>)(reinterpret_cast<T2*>(pb.A) + k * lda + k);
  const int tid = threadIdx.x, c = tid & (kGLeaf - 1), rr = tid / kGLeaf;
  const int lane = tid & 63, gbase = lane & ~(kGLeaf - 1), wave = tid >> 6;
  // the on-chip rows
  for (int e = tid; e < mc * kGLeaf; e += NT) {
    const int r = e / kGLeaf;
    cache[r][c] = c < wp ? A[(int64_t)r * lda + c] : zero2<T2>();
  }
  T2 reg[QR];
#pragma unroll
  for (int q = 0; q < QR; ++q) {
    const int r = LR + rr + RS * q;
    reg[q] = (r < mr && c < wp) ? A[(int64_t)r * lda + c] : zero2<T2>();
  }
  if (tid == 0) {
    snz = 0;
    sinv = zero2<T2>();
  }
  __syncthreads();
  for (int j = 0; j <= wp; ++j) {
    const bool upd = j > 0;
    const bool nz = snz != 0;
    const T2 inv = sinv;
    T2 u = zero2<T2>();
    if (upd && c < wp) u = urow[c];
    const bool act = c < wp && (upd ? c >= j - 1 : c == 0);
    double best = -1.0;
    int brow = INT_MAX;
    for (int base = j; base < mc; base += RS) {            // LDS rows
      const int r = base + rr;
      const bool v = r < mc;
      T2 x = (v && act) ? cache[r][c] : zero2<T2>();
      leaf_row(x, r, v, upd, nz, inv, u, c, j, gbase, best, brow);
      if (v && act) cache[r][c] = x;
    }
    if (mr > LR) {                                         // register rows
#pragma unroll
      for (int q = 0; q < QR; ++q) {
        const int r = LR + rr + RS * q;
        const bool v = r >= j && r < mr;
        T2 x = reg[q];
        leaf_row(x, r, v, upd, nz, inv, u, c, j, gbase, best, brow);
        if (v && act) reg[q] = x;
      }
    }
    for (int base = j > CR ? j : CR; base < m; base += RS * kGUnroll) {   // streamed rows
      T2 x[kGUnroll];
#pragma unroll
      for (int q = 0; q < kGUnroll; ++q) {
        const int r = base + rr + q * RS;
        x[q] = (act && r < m) ? A[(int64_t)r * lda + c] : zero2<T2>();
      }
#pragma unroll
      for (int q = 0; q < kGUnroll; ++q) {
        const int r = base + rr + q * RS;
        leaf_row(x[q], r, r < m, upd, nz, inv, u, c, j, gbase, best, brow);
        if (act && r < m) A[(int64_t)r * lda + c] = x[q];
      }
    }
    if (j == wp) break;
    // workgroup arg-max (every thread ends with the result)
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double ov = __shfl_xor(best, off);
      const int orow = __shfl_xor(brow, off);
      if (better(ov, orow, best, brow)) {
        best = ov;
        brow = orow;
      }
    }
    if (lane == 0) {
      sv[wave] = best;
      sr[wave] = brow;
    }
    __syncthreads();   // also orders this pass's stores before the swap below
    best = sv[0];
    brow = sr[0];
#pragma unroll
    for (int q = 1; q < NT / 64; ++q)
      if (better(sv[q], sr[q], best, brow)) {
        best = sv[q];
        brow = sr[q];
      }
    const int p = (brow == INT_MAX) ? j : brow;   // whole column NaN: keep row j
    // rows j and p wherever they live (LDS, a lane group's register slot,
    // memory): gather both, then write them swapped
    const int rels[2] = {j - LR, p - LR};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int r = t == 0 ? j : p;
      T2* dst = t == 0 ? rowa : rowb;
      if (r < LR) {
        if (tid < wp) dst[tid] = cache[r][tid];
      } else if (r < CR) {
        if (rr == rels[t] % RS && c < wp) {
          const int qs = rels[t] / RS;
#pragma unroll
          for (int q = 0; q < QR; ++q)
            if (q == qs) dst[c] = reg[q];
        }
      } else if (tid < wp) {
        dst[tid] = A[(int64_t)r * lda + tid];
      }
    }
    __syncthreads();
    if (p != j) {
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int r = t == 0 ? j : p;
        const T2* src = t == 0 ? rowb : rowa;
        if (r < LR) {
          if (tid < wp) cache[r][tid] = src[tid];
        } else if (r < CR) {
          if (rr == rels[t] % RS && c < wp) {
            const int qs = rels[t] / RS;
            const T2 v = src[c];
#pragma unroll
            for (int q = 0; q < QR; ++q)
              if (q == qs) reg[q] = v;
          }
        } else if (tid < wp) {
          A[(int64_t)r * lda + tid] = src[tid];
        }
      }
    }
    if (tid < wp) urow[tid] = rowb[tid];
    __syncthreads();
    if (tid == 0) {
      pb.ipiv[k + j] = (int)(k + p);
      const T2 piv = urow[j];
      const bool z = piv.x == 0 && piv.y == 0;
      snz = z ? 0 : 1;
      if (z) {
        sinv = zero2<T2>();
        atomicCAS(pb.info, 0, (int)(k + j + 1));
      } else {
        sinv = crecip(piv);
      }
    }
    __syncthreads();
  }
  // the on-chip rows back to the matrix (LDS rows were last written by other
  // lane groups in the final pass)
  __syncthreads();
  for (int e = tid; e < mc * kGLeaf; e += NT) {
    const int r = e / kGLeaf;
    if (c < wp) A[(int64_t)r * lda + c] = cache[r][c];
  }
#pragma unroll
  for (int q = 0; q < QR; ++q) {
    const int r = LR + rr + RS * q;
    if (r < mr && c < wp) A[(int64_t)r * lda + c] = reg[q];
  }
}

// ---- row swaps -------------------------------------------------------------
// Swaps k1 .. k2-1 (clamped to n) of each problem, in order, applied to
// columns [c0, c1) of its matrix (c1 clamped to n) or of its right-hand
// sides (rhs: c1 not clamped).  256 columns per workgroup.
struct GrpSwp {
  int64_t k1, k2, c0, c1;
  int rhs;
};

template <typename T2>
__global__ __launch_bounds__(256) void k_grp_laswp(const GrpProb* __restrict__ tab, GrpMap map, GrpSwp op) {
  const int p = grp_find(map, blockIdx.x);
  const GrpProb pb = tab[p];
  const int64_t c = op.c0 + (int64_t)(blockIdx.x - map.first[p]) * 256 + threadIdx.x;
  const int64_t c1 = op.rhs ? op.c1 : (op.c1 < pb.n ? op.c1 : pb.n);
  if (c >= c1) return;
  gptr<T2> A = (gptr<T2>)reinterpret_cast<T2*>(op.rhs ? pb.B : pb.A);
  const int64_t ld = op.rhs ? pb.ldb : pb.lda;
  const int64_t k2 = op.k2 < pb.n ? op.k2 : pb.n;
  for (int64_t i = op.k1; i < k2; ++i) {
    const int64_t r = pb.ipiv[i];
    if (r != i) {
      const T2 a = A[i * ld + c];
      A[i * ld + c] = A[r * ld + c];
      A[r * ld + c] = a;
    }
  }
}

// ---- triangular solves -----------------------------------------------------
// X := T^{-1} X for the diagonal block T at (r0, r0) of each problem's
// factors, of order w' = min(w, n - r0) <= kGTri: unit lower (forward) or
// upper (backward).  X: rows [r0, r0 + w'), columns [c0, c1) of the matrix
// (c1 clamped to n) or of the right-hand sides.  One row per lane; each wave
// carries kTriCols columns through the substitution together, four waves per
// chunk, kTriChunks chunks per workgroup (T is read into LDS once for them).
struct GrpTri {
  int64_t r0, c0, c1;
  int w, rhs;
};

constexpr int kTriCols = 8;
constexpr int kTriChunks = 1;   // column chunks per workgroup (more chunks: fewer loads of T, longer chains; 8 measured slower)
constexpr int kTriWgCols = 4 * kTriCols * kTriChunks;

template <typename T, bool UPPER>
__global__ __launch_bounds__(256) void k_grp_trsm(const GrpProb* __restrict__ tab, GrpMap map, GrpTri op) {
  typedef typename C2<T>::t T2;
  __shared__ T2 Ts[kGTri][kGTri + 1];
  __shared__ T2 dinv[kGTri];
  const int p = grp_find(map, blockIdx.x);
  const GrpProb pb = tab[p];
  const int wp = (int)(op.w < pb.n - op.r0 ? op.w : pb.n - op.r0);
  gcptr<T2> L = (gcptr<T2>)(reinterpret_cast<const T2*>(pb.A) + op.r0 * pb.lda + op.r0);
  for (int e = threadIdx.x; e < wp * wp; e += blockDim.x) Ts[e / wp][e % wp] = L[(e / wp) * pb.lda + e % wp];
  __syncthreads();
  if (UPPER && threadIdx.x < wp) dinv[threadIdx.x] = crecip(Ts[threadIdx.x][threadIdx.x]);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t c1 = op.rhs ? op.c1 : (op.c1 < pb.n ? op.c1 : pb.n);
  const int64_t ldx = op.rhs ? pb.ldb : pb.lda;
  gptr<T2> X = (gptr<T2>)(reinterpret_cast<T2*>(op.rhs ? pb.B : pb.A) + op.r0 * ldx);
  const int64_t cw = op.c0 + (int64_t)(blockIdx.x - map.first[p]) * kTriWgCols;
  for (int ch = 0; ch < kTriChunks; ++ch) {
  const int64_t cb = cw + (int64_t)ch * 4 * kTriCols + (threadIdx.x >> 6) * kTriCols;
  if (cw + (int64_t)ch * 4 * kTriCols >= c1) break;   // (uniform)
  T2 x[kTriCols];
#pragma unroll
  for (int q = 0; q < kTriCols; ++q) {
    x[q].x = 0;
    x[q].y = 0;
    if (cb + q < c1 && lane < wp) x[q] = X[lane * ldx + cb + q];
  }
  if (!UPPER) {
    for (int i = 0; i < wp - 1; ++i) {
      const T2 l = Ts[lane < wp ? lane : 0][i];
      const bool act = lane > i && lane < wp;
#pragma unroll
      for (int q = 0; q < kTriCols; ++q) {
        T2 xi;
        xi.x = __shfl(x[q].x, i);
        xi.y = __shfl(x[q].y, i);
        if (act) {
          x[q].x -= l.x * xi.x - l.y * xi.y;
          x[q].y -= l.x * xi.y + l.y * xi.x;
        }
      }
    }
  } else {
    for (int i = wp - 1; i >= 0; --i) {
      const T2 u = Ts[lane < wp ? lane : 0][i];
      const T2 di = dinv[i];
      const bool act = lane < i;
#pragma unroll
      for (int q = 0; q < kTriCols; ++q) {
        if (lane == i) x[q] = cmul(x[q], di);
        T2 xi;
        xi.x = __shfl(x[q].x, i);
        xi.y = __shfl(x[q].y, i);
        if (act) {
          x[q].x -= u.x * xi.x - u.y * xi.y;
          x[q].y -= u.x * xi.y + u.y * xi.x;
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kTriCols; ++q)
    if (cb + q < c1 && lane < wp) X[lane * ldx + cb + q] = x[q];
  }
}

// ---- products --------------------------------------------------------------
// C -= A B per problem: A = rows [m0, m1) x columns [k0, k1) of its matrix,
// B = rows [k0, k1) x columns [n0, n1) of its matrix (or right-hand sides),
// C = rows [m0, m1) x columns [n0, n1) of the same operand as B.  Ends are
// clamped to the problem's order (columns of right-hand sides are not).
// rhs: 0 = B and C from the matrix, 1 = from the right-hand sides,
// 2 = B from the auxiliary block X and C from the right-hand sides.
struct GrpGemm {
  int64_t m0, m1, n0, n1, k0, k1;
  int rhs;
};

struct GrpDims {
  int64_t M, N, K;
  int64_t ldb, ldc;
  const void* A;
  const void* B;
  void* C;
};

template <typename T2>
__device__ inline GrpDims grp_dims(const GrpProb& pb, const GrpGemm& op) {
  GrpDims d;
  d.M = clamp_extent(op.m1, pb.n, op.m0, true);
  d.N = clamp_extent(op.n1, pb.n, op.n0, !op.rhs);
  d.K = clamp_extent(op.k1, pb.n, op.k0, true);
  const T2* A = reinterpret_cast<const T2*>(pb.A);
  T2* Cb = reinterpret_cast<T2*>(op.rhs ? pb.B : pb.A);
  const T2* Bb = op.rhs == 2 ? reinterpret_cast<const T2*>(pb.X) : Cb;
  d.ldc = op.rhs ? pb.ldb : pb.lda;
  d.ldb = op.rhs == 2 ? pb.ldx : d.ldc;
  d.A = A + op.m0 * pb.lda + op.k0;
  d.B = Bb + op.k0 * d.ldb + op.n0;
  d.C = Cb + op.m0 * d.ldc + op.n0;
  return d;
}

template <typename T, int BM, int BN, int BK, int WM, int WN, bool G3>
__global__ __launch_bounds__(64 * WM * WN) void k_grp_gemm(const GrpProb* __restrict__ tab, GrpMap map, GrpGemm op) {
  typedef typename Cplx<T>::t T2;
  const int id = xcd_local(blockIdx.x, gridDim.x);
  const int p = grp_find(map, id);
  const GrpProb pb = tab[p];
  const GrpDims d = grp_dims<T2>(pb, op);
  const int nbm = (int)((d.M + BM - 1) / BM), nbn = (int)((d.N + BN - 1) / BN);
  // rows of 8 tiles walked column by column, as mg_tile_order's second step
  const int local = id - map.first[p];
  const int in_group = 8 * nbn, g = local / in_group, first_m = g * 8;
  const int gsize = min(nbm - first_m, 8);
  const int tm = first_m + (local % in_group) % gsize, tn = (local % in_group) / gsize;
  T2 alpha, beta;
  alpha.x = -1;
  alpha.y = 0;
  beta.x = 1;
  beta.y = 0;
  MemA a{d.A, pb.lda};
  cgemm_body<T, BM, BN, BK, WM, WN, MemA, G3>(d.M, d.N, d.K, alpha, a, d.B, d.ldb, beta, d.C, d.ldc, tm, tn);
}

// The grouped trailing update of complex-float factors on split-bf16 MFMA
// products (NS = 2: BF16x3), C -= A B.
template <int NS, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void k_grp_gemm_split(const GrpProb* __restrict__ tab, GrpMap map,
                                                                 GrpGemm op) {
  typedef mg_sg::Tile<NS, BM, BN, WM, WN> Tile;
  __shared__ typename Tile::S sm;
  const int id = xcd_local(blockIdx.x, gridDim.x);
  const int p = grp_find(map, id);
  const GrpProb pb = tab[p];
  const GrpDims d = grp_dims<float2>(pb, op);
  const int nbm = (int)((d.M + BM - 1) / BM), nbn = (int)((d.N + BN - 1) / BN);
  const int local = id - map.first[p];
  const int in_group = 8 * nbn, g = local / in_group, first_m = g * 8;
  const int gsize = min(nbm - first_m, 8);
  const int tm = first_m + (local % in_group) % gsize, tn = (local % in_group) / gsize;
  const int64_t i0 = (int64_t)tm * BM, j0 = (int64_t)tn * BN;
  const float2* A = reinterpret_cast<const float2*>(d.A);
  const float2* B = reinterpret_cast<const float2*>(d.B);
  float2* C = reinterpret_cast<float2*>(d.C);
  Tile t;
  // C - A B = -(A B - C): accumulators start at -C, the epilogue scales by -1
  t.init(C + i0 * d.ldc + j0, d.ldc, d.M - i0, d.N - j0, make_float2(-1.f, 0.f), true);
  t.mac(sm, A + i0 * pb.lda, pb.lda, B + j0, d.ldb, d.M - i0, d.N - j0, d.K);
  t.store(C + i0 * d.ldc + j0, d.ldc, d.M - i0, d.N - j0, make_float2(-1.f, 0.f), make_float2(1.f, 0.f), false);
}

// out[idx nrhs + c] = ||B[:, c]|| of every problem (caller's index idx),
// columns 64 at a time, the rows in 4 interleaved partial sums combined in a
// fixed order.
template <typename T2>
__global__ __launch_bounds__(256) void k_grp_colnorm(const GrpProb* __restrict__ tab, int nrhs, double* __restrict__ out) {
  __shared__ double part[4][64];
  const GrpProb pb = tab[blockIdx.y];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), ph = threadIdx.x >> 6;
  const T2* B = reinterpret_cast<const T2*>(pb.B);
  double acc = 0;
  if (c < nrhs)
    for (int64_t r = ph; r < pb.n; r += 4) {
      const T2 v = B[r * pb.ldb + c];
      acc += (double)v.x * v.x + (double)v.y * v.y;
    }
  part[ph][threadIdx.x & 63] = acc;
  __syncthreads();
  if (ph == 0 && c < nrhs)
    out[pb.idx * nrhs + c] = sqrt(((part[0][threadIdx.x] + part[1][threadIdx.x]) + part[2][threadIdx.x]) +
                                  part[3][threadIdx.x]);
}

// ---- BF16x3 trailing update from pre-split operands -------------------------
// Step k's update A22 -= L21 U12 (panel width nb = 32 T) reads its operands
// pre-split (psplit.h's format, BF16x3: four 16-byte units per row or column
// and group of 8 k): L21 (rows [k + nb, n) of columns [k, k + nb)) as A
// operand in SA[parity of the step], U12 (columns [k + nb, n) of rows [k, k +
// nb)) as B operand in SB.  Rows and columns past n - k - nb up to the next
// multiple of 128 hold zeros, so the update's 128 x 128 tiles (dma_update.h)
// read them unchecked.  L21 alternates between two buffers because step
// k + nb's panel is split on the look-ahead stream while step k's update
// still reads step k's.

__device__ inline int64_t pad128(int64_t v) { return (v + 127) / 128 * 128; }

// SA[par] of step k for every problem with rows below the step: thread t
// splits group t mod nb/8 of row t / (nb/8) (8 consecutive elements).
__global__ __launch_bounds__(256) void k_grp_pack_l(const GrpProb* __restrict__ tab, GrpMap map, int64_t k, int nb,
                                                    int par) {
  const int p = grp_find(map, blockIdx.x);
  const GrpProb pb = tab[p];
  const int kg = nb / 8;
  const int64_t r0 = k + nb, M = pb.n - r0;
  const int64_t t = (int64_t)(blockIdx.x - map.first[p]) * 256 + threadIdx.x;
  if (t >= pad128(M) * kg) return;
  const int64_t r = t / kg, g = t - r * kg;
  gcptr<float2> A = (gcptr<float2>)(reinterpret_cast<const float2*>(pb.A) + (r0 + r) * pb.lda + k + 8 * g);
  float2 v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = r < M ? (float2)A[e] : make_float2(0.f, 0.f);
  u32x4 u[4];
  mg_ps::split8<2>(v, u);
  gptr<u32x4> out = (gptr<u32x4>)(reinterpret_cast<u32x4*>(par ? pb.SA[1] : pb.SA[0]) + t * 4);   // (a select: no indexed copy of pb)
#pragma unroll
  for (int q = 0; q < 4; ++q) out[q] = u[q];
}

// SB of step k: thread t splits column t mod pad128(N) over group t / pad128(N)
// (consecutive threads read consecutive columns).
__global__ __launch_bounds__(256) void k_grp_pack_u(const GrpProb* __restrict__ tab, GrpMap map, int64_t k, int nb) {
  const int p = grp_find(map, blockIdx.x);
  const GrpProb pb = tab[p];
  const int kg = nb / 8;
  const int64_t c0 = k + nb, N = pb.n - c0, Np = pad128(N);
  const int64_t t = (int64_t)(blockIdx.x - map.first[p]) * 256 + threadIdx.x;
  if (t >= Np * kg) return;
  const int64_t g = t / Np, c = t - g * Np;
  gcptr<float2> B = (gcptr<float2>)(reinterpret_cast<const float2*>(pb.A) + (k + 8 * g) * pb.lda + c0 + c);
  float2 v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = c < N ? (float2)B[e * pb.lda] : make_float2(0.f, 0.f);
  u32x4 u[4];
  mg_ps::split8<2>(v, u);
  gptr<u32x4> out = (gptr<u32x4>)(reinterpret_cast<u32x4*>(pb.SB) + (c * kg + g) * 4);
#pragma unroll
  for (int q = 0; q < 4; ++q) out[q] = u[q];
}

// A22 -= L21 U12 of step k over columns [c0, c1) of the trailing matrix
// (counted from k + nb, c1 clamped to its order; c0 a multiple of 128):
// problem p's 128 x 128 tiles walked in rows of 8 tiles, column by column, as
// k_grp_gemm's.
template <int T>
__global__ __launch_bounds__(mg_dma::NT) void k_grp_update_dma(const GrpProb* __restrict__ tab, GrpMap map, int64_t k,
                                                              int64_t c0, int64_t c1, int par) {
  __shared__ mg_dma::Stage s0, s1;
  constexpr int KG = 4 * T;   // groups of 8 k: nb = 32 T
  const int id = xcd_local(blockIdx.x, gridDim.x);
  const int p = grp_find(map, id);
  const GrpProb pb = tab[p];
  const int64_t r0 = k + 8 * KG, M = pb.n - r0;
  const int64_t cend = c1 < M ? c1 : M;
  const int nbm = (int)((M + 127) / 128), nbn = (int)((cend - c0 + 127) / 128);
  const int local = id - map.first[p];
  const int in_group = 8 * nbn, g = local / in_group, first_m = g * 8;
  const int gsize = min(nbm - first_m, 8);
  const int tm = first_m + (local % in_group) % gsize, tn = (local % in_group) / gsize;
  const int64_t i0 = (int64_t)tm * 128, j0 = c0 + (int64_t)tn * 128;
  const u32x4* A = reinterpret_cast<const u32x4*>(par ? pb.SA[1] : pb.SA[0]) + i0 * KG * mg_dma::U;
  const u32x4* B = reinterpret_cast<const u32x4*>(pb.SB) + j0 * KG * mg_dma::U;
  float2* C = reinterpret_cast<float2*>(pb.A) + (r0 + i0) * pb.lda + r0 + j0;
  auto tile = [&](int t, mg_dma::Stage& st) { mg_dma::load_tile(st, A, B, KG, 4 * t); };
  // C read late (by DMA) where its 16-byte units are aligned and whole
  const int64_t Mt = M - i0, Nt = cend - j0;
  if ((reinterpret_cast<uintptr_t>(C) & 15) == 0 && pb.lda % 2 == 0 && (Nt >= 128 || Nt % 2 == 0))
    mg_dma::tile_update<T, true>(s0, s1, tile, C, pb.lda, Mt, Nt);
  else
    mg_dma::tile_update<T, false>(s0, s1, tile, C, pb.lda, Mt, Nt);
}

// ---- tournament-pivoting leaf (BF16x3 groups) --------------------------------
// The complex-float factors of a BF16x3 group only precondition an FP64
// refinement, so their leaves may pivot by tournament (communication-avoiding
// LU, Grigori, Demmel and Xiang, SIAM J. Matrix Anal. Appl. 32 (2011)): the
// leaf's rows [k, n) are cut into chunks of kTsRows; each chunk is eliminated
// with partial pivoting on its own and offers the w rows it pivoted on
// (k_grp_ts_local); one workgroup per problem eliminates the offered rows
// (their original values) with partial pivoting, which selects the w pivot
// rows in order and gives L11 \ U11, and moves the rows (k_grp_ts_final);
// every other row is then solved against U11 (k_grp_ts_apply).  Three
// launches of many short workgroups instead of one workgroup per problem that
// passes over every row once per column.  The pivots are those of partial
// pivoting on the offered rows, not LAPACK's; ipiv keeps LAPACK's meaning
// (row k + j swapped with row ipiv[k + j], in order).
constexpr int kTsRows = 256;     // rows per chunk
constexpr int kTsChunks = 64;    // chunks per problem at most: leaves of at most 16384 rows

// Workgroup arg-max of (v, r) over 256 threads (larger v first, then smaller
// r); every thread gets the result.
__device__ inline void ts_argmax(float& v, int& r, float* sv, int* sr) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const float ov = __shfl_xor(v, off);
    const int orr = __shfl_xor(r, off);
    if (ov > v || (ov == v && orr < r)) {
      v = ov;
      r = orr;
    }
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sv[wave] = v;
    sr[wave] = r;
  }
  __syncthreads();
  v = sv[0];
  r = sr[0];
#pragma unroll
  for (int q = 1; q < 4; ++q)
    if (sv[q] > v || (sv[q] == v && sr[q] < r)) {
      v = sv[q];
      r = sr[q];
    }
}

__device__ inline float cabs1(float2 a) { return fabsf(a.x) + fabsf(a.y); }

// Partial-pivoting elimination of the first wp columns of rows [0, rows) of
// t (row stride kGLeaf + 1), the rows' labels in lab swapped along; 256
// threads, row i held by thread i mod 256.  On return rows [0, wp) hold the
// pivot rows in order as L11 \ U11.
template <int RPT>
__device__ inline void ts_gepp(float2 (*t)[kGLeaf + 1], int* lab, int rows, int wp, float* sv, int* sr) {
  const int tid = threadIdx.x;
  for (int j = 0; j < wp; ++j) {
    float v = -1.f;
    int br = INT_MAX;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int i = tid + 256 * q;
      if (i >= j && i < rows && lab[i] >= 0) {
        const float a = cabs1(t[i][j]);
        if (a > v || (a == v && i < br)) {
          v = a;
          br = i;
        }
      }
    }
    ts_argmax(v, br, sv, sr);   // (its barrier also ends every thread's reads of column j)
    const int p = br == INT_MAX ? j : br;
    if (p != j) {
      if (tid < wp) {
        const float2 x = t[j][tid];
        t[j][tid] = t[p][tid];
        t[p][tid] = x;
      }
      if (tid == 0) {
        const int x = lab[j];
        lab[j] = lab[p];
        lab[p] = x;
      }
    }
    __syncthreads();
    const float2 piv = t[j][j];
    const bool nz = piv.x != 0.f || piv.y != 0.f;
    const float2 inv = nz ? crecip(piv) : make_float2(0.f, 0.f);
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int i = tid + 256 * q;
      if (i > j && i < rows) {
        float2 l = t[i][j];
        if (nz) l = cmul(l, inv);
        t[i][j] = l;
        for (int c = j + 1; c < wp; ++c) {
          const float2 u = t[j][c];
          t[i][c].x -= l.x * u.x - l.y * u.y;
          t[i][c].y -= l.x * u.y + l.y * u.x;
        }
      }
    }
    __syncthreads();
  }
}

// One workgroup per chunk of kTsRows rows of the leaf rows [k, n) x columns
// [k, k + wp): its wp pivot rows' indices (-1: past n) to cand[p][chunk][*]
// (p: the problem's place in the group).
__global__ __launch_bounds__(256) void k_grp_ts_local(const GrpProb* __restrict__ tab, GrpMap map, int64_t k, int w,
                                                      int* __restrict__ cand) {
  __shared__ float2 t[kTsRows][kGLeaf + 1];
  __shared__ int lab[kTsRows];
  __shared__ float sv[4];
  __shared__ int sr[4];
  const int p = grp_find(map, blockIdx.x);
  const GrpProb pb = tab[p];
  const int chunk = blockIdx.x - map.first[p];
  const int64_t m = pb.n - k, r0 = (int64_t)chunk * kTsRows;
  const int wp = (int)(w < m ? w : m);
  const int rows = (int)(m - r0 < kTsRows ? m - r0 : kTsRows);
  gcptr<float2> A = (gcptr<float2>)(reinterpret_cast<const float2*>(pb.A) + (k + r0) * pb.lda + k);
  for (int e = threadIdx.x; e < kTsRows * kGLeaf; e += 256) {
    const int r = e / kGLeaf, c = e % kGLeaf;
    t[r][c] = (r < rows && c < wp) ? (float2)A[(int64_t)r * pb.lda + c] : make_float2(0.f, 0.f);
  }
  lab[threadIdx.x] = threadIdx.x < rows ? (int)(k + r0) + threadIdx.x : -1;
  __syncthreads();
  ts_gepp<1>(t, lab, rows, wp, sv, sr);
  if (threadIdx.x < wp) cand[((int64_t)p * kTsChunks + chunk) * kGLeaf + threadIdx.x] = lab[threadIdx.x];
}

// One workgroup per problem: partial-pivoting elimination of the offered
// rows (original values), ipiv, info, the factored pivot rows to rows
// [k, k + wp) and the rows they displace to the pivot rows' places, all in
// columns [k, k + wp).
__global__ __launch_bounds__(256) void k_grp_ts_final(const GrpProb* __restrict__ tab, int64_t k, int w,
                                                      const int* __restrict__ cand) {
  constexpr int kCand = kTsChunks * kGLeaf;
  __shared__ float2 t[kCand][kGLeaf + 1];
  __shared__ int lab[kCand];
  __shared__ float2 top[kGLeaf][kGLeaf];
  __shared__ int mpos[2 * kGLeaf], msrc[2 * kGLeaf];
  __shared__ int nmove;
  __shared__ float sv[4];
  __shared__ int sr[4];
  const int p = blockIdx.x;
  const GrpProb pb = tab[p];
  const int64_t m = pb.n - k;
  const int wp = (int)(w < m ? w : m);
  const int nch = (int)((m + kTsRows - 1) / kTsRows);
  const int rows = nch * wp;
  gptr<float2> A = (gptr<float2>)reinterpret_cast<float2*>(pb.A);
  const int64_t lda = pb.lda;
  for (int e = threadIdx.x; e < rows * kGLeaf; e += 256) {
    const int r = e / kGLeaf, c = e % kGLeaf;
    const int g = cand[((int64_t)p * kTsChunks + r / wp) * kGLeaf + r % wp];
    t[r][c] = (g >= 0 && c < wp) ? A[(int64_t)g * lda + k + c] : make_float2(0.f, 0.f);
    if (c == 0) lab[r] = g;
  }
  for (int e = threadIdx.x; e < wp * kGLeaf; e += 256) {
    const int r = e / kGLeaf, c = e % kGLeaf;
    top[r][c] = c < wp ? A[(k + r) * lda + k + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  ts_gepp<kCand / 256>(t, lab, rows, wp, sv, sr);
  if (threadIdx.x == 0) {
    // LAPACK's swap sequence that brings lab[0 .. wp) to rows k .. k + wp - 1:
    // positions whose content changed, as (position, original row) pairs
    int cnt = 0;
    auto at = [&](int pos) {
      for (int i = 0; i < cnt; ++i)
        if (mpos[i] == pos) return msrc[i];
      return pos;
    };
    auto put = [&](int pos, int row) {
      for (int i = 0; i < cnt; ++i)
        if (mpos[i] == pos) {
          msrc[i] = row;
          return;
        }
      mpos[cnt] = pos;
      msrc[cnt] = row;
      ++cnt;
    };
    bool zero = false;
    for (int j = 0; j < wp; ++j) {
      const int target = lab[j], kj = (int)k + j;
      int q = target;
      for (int i = 0; i < cnt; ++i)
        if (msrc[i] == target) q = mpos[i];
      pb.ipiv[kj] = q;
      const int a = at(kj), b = at(q);
      put(kj, b);
      put(q, a);
      if (!zero && t[j][j].x == 0.f && t[j][j].y == 0.f) {
        zero = true;
        atomicCAS(pb.info, 0, kj + 1);
      }
    }
    nmove = cnt;
  }
  __syncthreads();
  // the factored pivot rows, then the displaced original rows of [k, k + wp)
  for (int e = threadIdx.x; e < wp * wp; e += 256) {
    const int r = e / wp, c = e % wp;
    A[(k + r) * lda + k + c] = t[r][c];
  }
  for (int e = threadIdx.x; e < nmove * wp; e += 256) {
    const int i = e / wp, c = e % wp;
    const int pos = mpos[i], src = msrc[i];
    if (pos >= k + wp && src >= k && src < k + wp) A[(int64_t)pos * lda + k + c] = top[src - k][c];
  }
}

// Rows [k + wp, n) x columns [k, k + wp) of every problem solved against
// U11 (row by row: the eliminations of the pivot rows), one row per thread.
__global__ __launch_bounds__(256) void k_grp_ts_apply(const GrpProb* __restrict__ tab, GrpMap map, int64_t k, int w) {
  __shared__ float2 u[kGLeaf][kGLeaf + 1];
  const int p = grp_find(map, blockIdx.x);
  const GrpProb pb = tab[p];
  const int64_t m = pb.n - k;
  const int wp = (int)(w < m ? w : m);
  gptr<float2> A = (gptr<float2>)reinterpret_cast<float2*>(pb.A);
  const int64_t lda = pb.lda;
  for (int e = threadIdx.x; e < wp * kGLeaf; e += 256) {
    const int r = e / kGLeaf, c = e % kGLeaf;
    u[r][c] = c < wp ? A[(k + r) * lda + k + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const int64_t r = k + wp + (int64_t)(blockIdx.x - map.first[p]) * 256 + threadIdx.x;
  if (r >= pb.n) return;
  float2 x[kGLeaf];
  gptr<float2> row = A + r * lda + k;
#pragma unroll
  for (int c = 0; c < kGLeaf; ++c) x[c] = c < wp ? row[c] : make_float2(0.f, 0.f);
#pragma unroll
  for (int j = 0; j < kGLeaf; ++j) {
    if (j < wp) {
      const float2 piv = u[j][j];
      float2 l = x[j];
      if (piv.x != 0.f || piv.y != 0.f) l = cmul(l, crecip(piv));
      x[j] = l;
#pragma unroll
      for (int c = j + 1; c < kGLeaf; ++c) {
        const float2 v = u[j][c];
        x[c].x -= l.x * v.x - l.y * v.y;
        x[c].y -= l.x * v.y + l.y * v.x;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < kGLeaf; ++c)
    if (c < wp) row[c] = x[c];
}

// ---- BF16x3 groups: solves from inverted diagonal blocks ---------------------
// The factors of a BF16x3 group precondition an FP64 refinement, so their
// solves may run on GEMMs only: the 64 x 64 diagonal blocks of L (unit) and U
// are inverted once per factorisation (Dinv: per block, L^-1 then U^-1, 64 x
// 64 row-major, zero past the block's order), and each sweep is one dataflow
// launch over the blocks of every problem (k_grp_sweep): the workgroup of
// block k (64 rows, 32 right-hand sides) forms W_k - sum_j T_kj X_j over the
// blocks j it depends on (forward: j < k with T = L; backward: j > k with T
// = U) as soon as each X_j is published, then X_k = D_k^-1 (that), publishes
// it and signals.  Products on split-bf16 MFMA (BF16x3), 64 x 32 tiles of
// 128 threads.
constexpr int kSB = 64;   // solve block
typedef mg_sg::Tile<2, kSB, 32, 2, 1> SolveTile;

// perm from getrf's swap sequence, one workgroup per problem (the swaps
// composed by one thread in LDS for n <= kPermLds, in memory beyond).
constexpr int kPermLds = 16384;
__global__ __launch_bounds__(256) void k_grp_perm(const GrpProb* __restrict__ tab) {
  __shared__ int sp[kPermLds], sv[kPermLds];
  const GrpProb pb = tab[blockIdx.x];
  const int n = (int)pb.n;
  gptr<int> perm = (gptr<int>)pb.perm;
  gcptr<int> ipiv = (gcptr<int>)pb.ipiv;
  if (n <= kPermLds) {
    for (int r = threadIdx.x; r < n; r += 256) {
      sp[r] = r;
      sv[r] = ipiv[r];
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 0; i < n; ++i) {
        const int j = sv[i], t = sp[i];
        sp[i] = sp[j];
        sp[j] = t;
      }
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += 256) perm[r] = sp[r];
  } else {
    for (int r = threadIdx.x; r < n; r += 256) perm[r] = r;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 0; i < n; ++i) {
        const int j = ipiv[i], t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
      }
  }
}

// Inverse of diagonal block b (order h = min(64, n - 64 b)) of L (unit
// lower; workgroup 2 b of the problem) or U (upper; 2 b + 1), by columns:
// lane c substitutes column c of the identity.  One wave per workgroup.
__global__ __launch_bounds__(64) void k_grp_trinv(const GrpProb* __restrict__ tab, GrpMap map) {
  __shared__ float2 t[kSB][kSB + 1], x[kSB][kSB + 1];
  const int p = grp_find(map, blockIdx.x);
  const GrpProb pb = tab[p];
  const int local = blockIdx.x - map.first[p], b = local >> 1;
  const bool upper = local & 1;
  const int64_t r0 = (int64_t)b * kSB;
  const int h = (int)(pb.n - r0 < kSB ? pb.n - r0 : kSB);
  gcptr<float2> A = (gcptr<float2>)(reinterpret_cast<const float2*>(pb.A) + r0 * pb.lda + r0);
  const int c = threadIdx.x;
  for (int r = 0; r < h; ++r) t[r][c] = c < h ? (float2)A[r * pb.lda + c] : make_float2(0.f, 0.f);
  __syncthreads();
  for (int r = 0; r < kSB; ++r) x[r][c] = make_float2(0.f, 0.f);
  if (c < h) {
    if (!upper) {
      x[c][c] = make_float2(1.f, 0.f);
      for (int r = c + 1; r < h; ++r) {
        float2 acc = make_float2(0.f, 0.f);
        for (int j = c; j < r; ++j) {
          const float2 l = t[r][j], v = x[j][c];
          acc.x -= l.x * v.x - l.y * v.y;
          acc.y -= l.x * v.y + l.y * v.x;
        }
        x[r][c] = acc;
      }
    } else {
      x[c][c] = crecip(t[c][c]);
      for (int r = c - 1; r >= 0; --r) {
        float2 acc = make_float2(0.f, 0.f);
        for (int j = r + 1; j <= c; ++j) {
          const float2 u = t[r][j], v = x[j][c];
          acc.x -= u.x * v.x - u.y * v.y;
          acc.y -= u.x * v.y + u.y * v.x;
        }
        x[r][c] = cmul(acc, crecip(t[r][r]));
      }
    }
  }
  __syncthreads();
  gptr<float2> D = (gptr<float2>)(reinterpret_cast<float2*>(pb.Dinv) + ((int64_t)2 * b + (upper ? 1 : 0)) * kSB * kSB);
  for (int r = 0; r < kSB; ++r) D[r * kSB + c] = x[r][c];
}

// W1 = P B (rows gathered through perm), 256 entries per workgroup.
__global__ __launch_bounds__(256) void k_grp_gather(const GrpProb* __restrict__ tab, GrpMap map, int nrhs) {
  const int p = grp_find(map, blockIdx.x);
  const GrpProb pb = tab[p];
  const int64_t e = (int64_t)(blockIdx.x - map.first[p]) * 256 + threadIdx.x;
  if (e >= pb.n * nrhs) return;
  const int64_t r = e / nrhs, c = e - r * nrhs;
  gcptr<float2> B = (gcptr<float2>)reinterpret_cast<const float2*>(pb.B);
  gptr<float2> W = (gptr<float2>)reinterpret_cast<float2*>(pb.W1);
  W[r * nrhs + c] = B[(int64_t)((gcptr<int>)pb.perm)[r] * pb.ldb + c];
}

// Tickets of a sweep in dependency order: level s holds block s (forward) or
// block nblk - 1 - s (backward) of every problem with more than s blocks
// (a prefix of the group: orders decrease), nct column tiles each; level s
// starts at ticket lev[s].  A workgroup only waits for blocks of lower levels,
// whose tickets were taken before its own, so the sweep needs no
// co-residency.
// ctr (zero on entry): ctr[0] hands out tickets, ctr[1 + coff_p + k] counts
// the finished column tiles of problem p's block k (coff_p = blocks of the
// problems before p); next: the following sweep's counters, cleared here.
// Published blocks are written with agent-scope stores and read with
// agent-scope loads (coherent across XCDs), each wave's stores complete
// before the barrier that precedes the signal.  A wait past the spin bound
// publishes NaN, so every workgroup ends and the refinement sees the failure.
constexpr long long kGrpSweepSpin = 1LL << 22;

template <bool FWD>
__global__ __launch_bounds__(128) void k_grp_sweep(const GrpProb* __restrict__ tab, const int* __restrict__ lev,
                                                   int nlev, const int* __restrict__ coff, int nrhs, int* ctr,
                                                   int* next, int cap) {
  __shared__ SolveTile::S sm;
  __shared__ int s_ticket, s_ok;
  if (blockIdx.x == 0)
    for (int e = threadIdx.x; e < cap; e += blockDim.x)
      __hip_atomic_store(next + e, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (threadIdx.x == 0) s_ticket = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int ticket = s_ticket;
  int lo = 0, hi = nlev;   // the level: the last s with lev[s] <= ticket
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (lev[mid] <= ticket) lo = mid;
    else hi = mid;
  }
  const int lvl = lo;
  const int nct = (nrhs + 31) / 32;
  const int p = (ticket - lev[lvl]) / nct, ct = (ticket - lev[lvl]) % nct;
  const GrpProb pb = tab[p];
  const int nblk = (int)((pb.n + kSB - 1) / kSB);
  const int k = FWD ? lvl : nblk - 1 - lvl;
  int* done = ctr + 1 + coff[p];
  const int64_t c0 = (int64_t)ct * 32, N = nrhs - c0 < 32 ? nrhs - c0 : 32;
  const int64_t rk = (int64_t)k * kSB, hk = pb.n - rk < kSB ? pb.n - rk : kSB;
  float2* Wk = reinterpret_cast<float2*>(FWD ? pb.W1 : pb.W2) + rk * nrhs + c0;   // this block's input, in place
  float2* X = reinterpret_cast<float2*>(FWD ? pb.W2 : pb.B);                       // the published blocks
  const int64_t ldx = FWD ? nrhs : pb.ldb;
  SolveTile t;
  t.init(Wk, nrhs, hk, N, make_float2(-1.f, 0.f), true);   // -W_k
  bool ok = true;
  for (int q = 0; q < (FWD ? k : nblk - 1 - k); ++q) {
    const int j = FWD ? q : nblk - 1 - q;
    if (threadIdx.x == 0) {
      long long spins = 0;
      int st = 1;
      while (__hip_atomic_load(&done[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nct) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kGrpSweepSpin) {
          st = 0;
          break;
        }
      }
      s_ok = st;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (!s_ok) {
      ok = false;
      break;
    }
    // += T_kj X_j (rows of block k, columns of block j; X_j agent-scope loads)
    const int64_t rj = (int64_t)j * kSB, hj = pb.n - rj < kSB ? pb.n - rj : kSB;
    const float2* T = reinterpret_cast<const float2*>(pb.A) + rk * pb.lda + rj;
    const float2* Xj = X + rj * ldx + c0;
    t.template fetch<true>(T, pb.lda, Xj, ldx, hk, N, hj, 0);
    for (int64_t k0 = 0; k0 < hj; k0 += mg_sg::BK) {
      __syncthreads();
      t.stash(sm);
      __syncthreads();
      if (k0 + mg_sg::BK < hj) t.template fetch<true>(T, pb.lda, Xj, ldx, hk, N, hj, k0 + mg_sg::BK);
      t.compute(sm);
    }
  }
  float2* Xk = X + rk * ldx + c0;
  if (ok) {
    // W_k' = -acc to this block's input rows (read back below by this workgroup only)
    t.store(Wk, nrhs, hk, N, make_float2(-1.f, 0.f), make_float2(1.f, 0.f), false);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const float2* Dk = reinterpret_cast<const float2*>(pb.Dinv) + ((int64_t)2 * k + (FWD ? 0 : 1)) * kSB * kSB;
    SolveTile d;
    d.init(nullptr, 0, 0, 0, make_float2(0.f, 0.f), false);
    d.mac(sm, Dk, kSB, Wk, nrhs, hk, N, hk);
    d.template store<true>(Xk, ldx, hk, N, make_float2(1.f, 0.f), make_float2(0.f, 0.f), false);
  } else {
    const float nan = __builtin_nanf("");
    for (int e = threadIdx.x; e < hk * N; e += blockDim.x)
      mg_sg::coherent_store(&Xk[(e / N) * ldx + e % N], make_float2(nan, nan));
  }
  // every wave's stores complete before the barrier after which the block's
  // counter is bumped (a workgroup fence emits no such wait on gfx950)
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_fetch_add(&done[k], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_grp_clear(int* c, int cap) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < cap; e += gridDim.x * blockDim.x)
    __hip_atomic_store(c + e, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- FP64 residual of the refinement: B -= A X ------------------------------
// Complex128 A (n x n) and X (n x nrhs, the auxiliary block), B the
// right-hand sides.  One pass over A at streaming rate: a workgroup of four
// waves owns 64 RF rows (16 RF per wave: RF 16-row MFMA fragments) and 32
// right-hand sides and walks k in chunks of 4 KS (KS 16x16x4 FP64 MFMA
// steps); each lane's A and X fragments of the next chunk are loaded into
// registers while the current chunk is multiplied, straight from memory
// (every element of A is read once; X, n x 32, through the caches).
// Complex products in the three-product form (t1 = Ar Xr, t2 = Ai Xi, t3 =
// (Ar + Ai)(Xr + Xi)).  Rows, k and columns past the problem's extents are
// loaded from clamped addresses and zeroed or not stored.
// Measured on the 256-matrix ensemble (96 GB of complex128, 32 columns):
// RF 2, KS 2 at two waves per SIMD 36 ms; RF 1 with KS 4 (three waves per
// SIMD) 41 ms, with KS 2 44 ms; the rows staged through LDS by DMA in
// 512-byte segments (three stages, one workgroup per CU) 59 ms; the FP64
// DMA-staged GEMM tile of the Krylov product at 128 x 32 (three stages of
// 40 KB, one workgroup per CU) 37 ms.
constexpr int kResCols = 32;

template <int RF, int KS>
struct ResFrag {
  double2 a[RF][KS];   // [row fragment][k-step]
  double2 x[2][KS];    // [column fragment][k-step]
};

template <int RF, int KS, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_grp_residual(
    const GrpProb* __restrict__ tab, GrpMap map, int nrhs) {
  const int id = xcd_local(blockIdx.x, gridDim.x);
  const int p = grp_find(map, id);
  const GrpProb pb = tab[p];
  const int64_t n = pb.n;
  const int local = id - map.first[p];
  const int nct = (nrhs + kResCols - 1) / kResCols;
  const int rb = local / nct, ct = local - rb * nct;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  const int64_t r0 = (int64_t)rb * (64 * RF) + wave * (16 * RF);
  const int c0 = ct * kResCols;
  gcptr<double2> A = (gcptr<double2>)reinterpret_cast<const double2*>(pb.A);
  gcptr<double2> X = (gcptr<double2>)reinterpret_cast<const double2*>(pb.X);
  // this lane's A entries: rows r0 + 16 f + li (clamped), k = k0 + 4 s + lk;
  // X entries: k = k0 + 4 s + lk, columns c0 + 16 g + li (clamped)
  gcptr<double2> pa[RF];
  gcptr<double2> px[2];
#pragma unroll
  for (int f = 0; f < RF; ++f) {
    const int64_t r = r0 + 16 * f + li;
    pa[f] = A + (r < n ? r : n - 1) * pb.lda + lk;
  }
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const int64_t c = c0 + 16 * g + li;
    px[g] = X + (int64_t)lk * pb.ldx + (c < nrhs ? c : nrhs - 1);
  }
  const int64_t xstep = 4 * pb.ldx;
  typedef ResFrag<RF, KS> Frag;
  // full chunks: unpredicated loads at k0 (pointers advanced by the caller)
  auto fetch = [&](Frag& fr) {
#pragma unroll
    for (int s = 0; s < KS; ++s) {
#pragma unroll
      for (int f = 0; f < RF; ++f) fr.a[f][s] = pa[f][4 * s];
#pragma unroll
      for (int g = 0; g < 2; ++g) fr.x[g][s] = px[g][s * xstep];
    }
  };
  constexpr int64_t KC = 4 * KS;
  auto advance = [&]() {
#pragma unroll
    for (int f = 0; f < RF; ++f) pa[f] += KC;
#pragma unroll
    for (int g = 0; g < 2; ++g) px[g] += KC * pb.ldx;
  };
  v4d t1[RF][2], t2[RF][2], t3[RF][2];
#pragma unroll
  for (int f = 0; f < RF; ++f)
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int r = 0; r < 4; ++r) t1[f][g][r] = t2[f][g][r] = t3[f][g][r] = 0.0;
  auto step = [&](const Frag& fr) {
#pragma unroll
    for (int s = 0; s < KS; ++s) {
#pragma unroll
      for (int f = 0; f < RF; ++f) {
        const double ar = fr.a[f][s].x, ai = fr.a[f][s].y, as = ar + ai;
#pragma unroll
        for (int g = 0; g < 2; ++g) {
          const double xr = fr.x[g][s].x, xi = fr.x[g][s].y, xs = xr + xi;
          t1[f][g] = mg_mfma(ar, xr, t1[f][g]);
          t2[f][g] = mg_mfma(ai, xi, t2[f][g]);
          t3[f][g] = mg_mfma(as, xs, t3[f][g]);
        }
      }
    }
  };
  const int64_t nfull = n / KC;   // whole chunks; the rest (n mod KC) below
  Frag b0, b1;
  if (nfull > 0) {
    fetch(b0);
    advance();
  }
  // steady state with unconditional fetches (the compiler's waits then only
  // cover the chunk being multiplied), then the last one or two chunks
  int64_t c = 0;
  for (; c + 2 < nfull; c += 2) {
    fetch(b1);
    advance();
    step(b0);
    fetch(b0);
    advance();
    step(b1);
  }
  if (c + 1 < nfull) {
    fetch(b1);
    advance();
    step(b0);
    step(b1);
  } else if (c < nfull) {
    step(b0);
  }
  if (nfull * KC < n) {   // the last partial chunk, k past n zeroed
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const bool ok = nfull * KC + 4 * s + lk < n;
#pragma unroll
      for (int f = 0; f < RF; ++f) b0.a[f][s] = ok ? (double2)pa[f][4 * s] : make_double2(0.0, 0.0);
#pragma unroll
      for (int g = 0; g < 2; ++g) b0.x[g][s] = ok ? (double2)px[g][s * xstep] : make_double2(0.0, 0.0);
    }
    step(b0);
  }
  // accumulator item r of lane l: row l / 16 + 4 r, column l % 16 (FP64 MFMA)
  gptr<double2> B = (gptr<double2>)reinterpret_cast<double2*>(pb.B);
#pragma unroll
  for (int f = 0; f < RF; ++f)
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = r0 + 16 * f + lk + 4 * r;
        const int col = c0 + 16 * g + li;
        if (row < n && col < nrhs) {
          double2 v = B[row * pb.ldb + col];
          v.x -= t1[f][g][r] - t2[f][g][r];
          v.y -= t3[f][g][r] - t1[f][g][r] - t2[f][g][r];
          B[row * pb.ldb + col] = v;
        }
      }
}

__global__ void k_grp_zero_info(const GrpProb* __restrict__ tab, int P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < P) *tab[p].info = 0;
}

inline int64_t split_half(int64_t w) {
  // left part: half, rounded up to a multiple of 16 (keeps product tiles whole)
  int64_t h = ((w / 2) + 15) / 16 * 16;
  if (h >= w) h = w / 2;
  return h > 0 ? h : 1;
}

}  // namespace

namespace mg {

// A group of problems on one device: sizes sorted by decreasing order, the
// device table, the tunables and an optional captured graph of the factor
// and solve sequence.
struct Group {
  int prec = MG_C128;
  int P = 0;
  std::vector<int64_t> n;     // sorted, decreasing
  std::vector<GrpProb> host;  // the table, same order
  std::vector<int> order;     // order[i]: caller index of sorted entry i
  GrpProb* dtab = nullptr;
  int64_t nrhs = 0;
  GrpCfg cfg;
  hipGraph_t graph[2] = {nullptr, nullptr};        // captured factor (0) and solve (1) sequences
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipStream_t s = nullptr;    // the stream launches go to (the call's, or s2 for look-ahead panels)
  hipStream_t s2 = nullptr;   // look-ahead panel stream (default priority)
  hipStream_t scap = nullptr; // capture stream (default priority)
  hipEvent_t evA = nullptr, evB = nullptr;
  int64_t launches = 0;
  void* split = nullptr;      // BF16x3: the pre-split operand buffers of every problem
  int split_nb = 0;           // the panel width they were laid out for
  int* cand = nullptr;        // BF16x3: the rows each chunk of a tournament leaf offers
  void* solvebuf = nullptr;   // BF16x3 with right-hand sides: perm, Dinv, W1, W2 of every problem
  int* sweep_lev = nullptr;   // BF16x3 solves: first ticket of each sweep level (device, in solvebuf)
  int* sweep_coff = nullptr;  // BF16x3 solves: each problem's first block counter (device, in solvebuf)
  int* sweep_ctr = nullptr;   // BF16x3 solves: two counter sets of sweep_cap ints (device, in solvebuf)
  int sweep_nlev = 0, sweep_cap = 0, sweep_tickets = 0;

  // problems with rows at or beyond row k (a prefix: orders decrease)
  int active(int64_t k) const {
    int a = 0;
    while (a < P && n[a] > k) ++a;
    return a;
  }

  // tile table of a launch: problem p gets cnt(p) workgroups
  template <typename F>
  bool make_map(int np, F cnt, GrpMap& m, int& total) const {
    m.np = np;
    int64_t acc = 0;
    for (int p = 0; p < np; ++p) {
      m.first[p] = (int)acc;
      acc += cnt(p);
    }
    m.first[np] = (int)acc;
    for (int p = np + 1; p <= kMaxGroup; ++p) m.first[p] = (int)acc;
    total = (int)acc;
    return acc > 0 && acc < INT_MAX;
  }
};

namespace {

template <typename T>
int grp_getf2(Group& g, int64_t k, int w) {
  const int np = g.active(k);
  if (np == 0) return MG_OK;
  switch (g.cfg.lthreads) {
    case 1024: hipLaunchKernelGGL((k_grp_getf2<T, 1024>), dim3(np), dim3(1024), 0, g.s, g.dtab, k, w); break;
    case 512: hipLaunchKernelGGL((k_grp_getf2<T, 512>), dim3(np), dim3(512), 0, g.s, g.dtab, k, w); break;
    default: hipLaunchKernelGGL((k_grp_getf2<T, 256>), dim3(np), dim3(256), 0, g.s, g.dtab, k, w); break;
  }
  MG_HIP(hipGetLastError());
  ++g.launches;
  return MG_OK;
}

// The BF16x3 trailing update from pre-split operands applies to panel widths
// that are whole 128-column tiles with an even count of 32-deep k-tiles.
bool dma_split(const Group& g) { return g.prec == MG_BF16X3 && (g.cfg.nb == 128 || g.cfg.nb == 256 || g.cfg.nb == 384); }

// The pre-split operand buffers of a BF16x3 group: per problem SA[0], SA[1]
// and SB of pad128(n) rows or columns of nb / 8 groups of four 16-byte units,
// in one allocation, laid out for the current panel width and kept while it
// is not wider.  Called before a factorisation or its capture (never inside
// one); a new allocation waits for the device first, as work in flight may
// still read the old one.
int ensure_split(Group& g) {
  if (g.prec != MG_BF16X3 || g.P == 0) return MG_OK;
  if (!g.cand && hipMalloc(&g.cand, sizeof(int) * g.P * kTsChunks * kGLeaf) != hipSuccess) {
    g.cand = nullptr;
    set_error("grouped LU: tournament candidate table not allocated");
    return MG_ERR_NOMEM;
  }
  if (g.nrhs > 0 && !g.solvebuf) {   // the solve's buffers (sizes fixed at creation)
    int64_t total = 0;
    auto sizes = [&](int p, int64_t& perm, int64_t& dinv, int64_t& w) {
      const int64_t nblk = (g.n[p] + kSB - 1) / kSB;
      perm = (g.n[p] * (int64_t)sizeof(int) + 255) / 256 * 256;
      dinv = 2 * nblk * kSB * kSB * (int64_t)sizeof(float2);
      w = (g.n[p] * g.nrhs * (int64_t)sizeof(float2) + 255) / 256 * 256;
    };
    for (int p = 0; p < g.P; ++p) {
      int64_t a, b, c;
      sizes(p, a, b, c);
      total += a + b + 2 * c;
    }
    // the sweeps' ticket levels, counter offsets and counters
    const int nct = (int)((g.nrhs + 31) / 32);
    const int nlev = (int)((g.n[0] + kSB - 1) / kSB);
    std::vector<int> lev(nlev + 1, 0), coff(g.P, 0);
    int64_t blocks = 0;
    for (int p = 0; p < g.P; ++p) {
      coff[p] = (int)blocks;
      blocks += (g.n[p] + kSB - 1) / kSB;
    }
    for (int l = 0, np = g.P; l < nlev; ++l) {
      while (np > 0 && (g.n[np - 1] + kSB - 1) / kSB <= l) --np;
      lev[l + 1] = lev[l] + np * nct;
    }
    if (blocks + 1 > INT_MAX / 4 || (int64_t)lev[nlev] > INT_MAX / 2) {
      set_error("grouped LU: too many solve blocks");
      return MG_ERR_ARG;
    }
    const int cap = (int)blocks + 1;
    const int64_t tabs = ((int64_t)(nlev + 1 + g.P + 2 * cap) * (int64_t)sizeof(int) + 255) / 256 * 256;
    total += tabs;
    if (hipMalloc(&g.solvebuf, total) != hipSuccess) {
      g.solvebuf = nullptr;
      set_error("grouped LU: solve buffers (%lld bytes) not allocated", (long long)total);
      return MG_ERR_NOMEM;
    }
    char* q = static_cast<char*>(g.solvebuf);
    for (int p = 0; p < g.P; ++p) {
      int64_t a, b, c;
      sizes(p, a, b, c);
      g.host[p].perm = reinterpret_cast<int*>(q);
      g.host[p].Dinv = q + a;
      g.host[p].W1 = q + a + b;
      g.host[p].W2 = q + a + b + c;
      q += a + b + 2 * c;
    }
    g.sweep_lev = reinterpret_cast<int*>(q);
    g.sweep_coff = g.sweep_lev + nlev + 1;
    g.sweep_ctr = g.sweep_coff + g.P;
    g.sweep_nlev = nlev;
    g.sweep_cap = cap;
    MG_HIP(hipMemcpy(g.sweep_lev, lev.data(), sizeof(int) * (nlev + 1), hipMemcpyHostToDevice));
    MG_HIP(hipMemcpy(g.sweep_coff, coff.data(), sizeof(int) * g.P, hipMemcpyHostToDevice));
    MG_HIP(hipMemset(g.sweep_ctr, 0, sizeof(int) * 2 * cap));
    MG_HIP(hipMemcpy(g.dtab, g.host.data(), sizeof(GrpProb) * g.P, hipMemcpyHostToDevice));
    g.sweep_tickets = lev[nlev];
  }
  if (!dma_split(g)) return MG_OK;
  if (g.split && g.split_nb >= g.cfg.nb) return MG_OK;
  if (g.split) {
    MG_HIP(hipDeviceSynchronize());
    MG_HIP(hipFree(g.split));
    g.split = nullptr;
    g.split_nb = 0;
  }
  const int64_t per = (int64_t)g.cfg.nb / 8 * 4 * 16;   // bytes per row or column
  int64_t total = 0;
  for (int p = 0; p < g.P; ++p) total += 3 * ((g.n[p] + 127) / 128 * 128) * per;
  if (hipMalloc(&g.split, total) != hipSuccess) {
    g.split = nullptr;
    set_error("grouped LU: pre-split operand buffers (%lld bytes) not allocated", (long long)total);
    return MG_ERR_NOMEM;
  }
  char* b = static_cast<char*>(g.split);
  for (int p = 0; p < g.P; ++p) {
    const int64_t sz = ((g.n[p] + 127) / 128 * 128) * per;
    g.host[p].SA[0] = b;
    g.host[p].SA[1] = b + sz;
    g.host[p].SB = b + 2 * sz;
    b += 3 * sz;
  }
  MG_HIP(hipMemcpy(g.dtab, g.host.data(), sizeof(GrpProb) * g.P, hipMemcpyHostToDevice));
  g.split_nb = g.cfg.nb;
  return MG_OK;
}

// Tournament-pivoting leaves apply to BF16x3 groups whose leaf has at most
// kTsRows * kTsChunks rows.
bool ts_leaf(const Group& g, int64_t k) {
  return g.prec == MG_BF16X3 && g.cfg.pivot == 1 && g.cand && g.n[0] - k <= (int64_t)kTsRows * kTsChunks;
}

// Tournament-pivoting leaf over columns [k, k + w) of every problem.
int grp_tslu(Group& g, int64_t k, int w) {
  const int np = g.active(k);
  if (np == 0) return MG_OK;
  GrpMap m;
  int total;
  if (!g.make_map(np, [&](int p) -> int64_t { return (g.n[p] - k + kTsRows - 1) / kTsRows; }, m, total))
    return MG_OK;
  hipLaunchKernelGGL(k_grp_ts_local, dim3(total), dim3(256), 0, g.s, g.dtab, m, k, w, g.cand);
  MG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_grp_ts_final, dim3(np), dim3(256), 0, g.s, g.dtab, k, w, g.cand);
  MG_HIP(hipGetLastError());
  g.launches += 2;
  if (g.make_map(np, [&](int p) -> int64_t {
        const int64_t m = g.n[p] - k, below = m - std::min<int64_t>(w, m);
        return (below + 255) / 256;
      }, m, total)) {
    hipLaunchKernelGGL(k_grp_ts_apply, dim3(total), dim3(256), 0, g.s, g.dtab, m, k, w);
    MG_HIP(hipGetLastError());
    ++g.launches;
  }
  return MG_OK;
}

// Split L21 (lower) into SA[par] or U12 into SB for step k.
int grp_pack(Group& g, int64_t k, bool lower, int par) {
  const int nb = g.cfg.nb;
  const int64_t kn = k + nb;
  const int np = g.active(kn);
  if (np == 0) return MG_OK;
  GrpMap m;
  int total;
  if (!g.make_map(np, [&](int p) -> int64_t { return (((g.n[p] - kn) + 127) / 128 * 128 * (nb / 8) + 255) / 256; },
                  m, total))
    return MG_OK;
  if (lower) hipLaunchKernelGGL(k_grp_pack_l, dim3(total), dim3(256), 0, g.s, g.dtab, m, k, nb, par);
  else hipLaunchKernelGGL(k_grp_pack_u, dim3(total), dim3(256), 0, g.s, g.dtab, m, k, nb);
  MG_HIP(hipGetLastError());
  ++g.launches;
  return MG_OK;
}

// A22 -= L21 U12 of step k from the pre-split operands, columns [c0, c1) of
// the trailing matrix (counted from k + nb).
int grp_update_dma(Group& g, int64_t k, int64_t c0, int64_t c1, int par) {
  const int nb = g.cfg.nb;
  const int64_t kn = k + nb;
  const int np = g.active(kn);
  if (np == 0) return MG_OK;
  GrpMap m;
  int total;
  if (!g.make_map(np, [&](int p) -> int64_t {
        const int64_t M = g.n[p] - kn, cend = std::min(c1, M);
        if (M <= 0 || cend <= c0) return 0;
        return ((M + 127) / 128) * ((cend - c0 + 127) / 128);
      }, m, total))
    return MG_OK;
  switch (nb) {
    case 128: hipLaunchKernelGGL(k_grp_update_dma<4>, dim3(total), dim3(mg_dma::NT), 0, g.s, g.dtab, m, k, c0, c1, par); break;
    case 256: hipLaunchKernelGGL(k_grp_update_dma<8>, dim3(total), dim3(mg_dma::NT), 0, g.s, g.dtab, m, k, c0, c1, par); break;
    case 384: hipLaunchKernelGGL(k_grp_update_dma<12>, dim3(total), dim3(mg_dma::NT), 0, g.s, g.dtab, m, k, c0, c1, par); break;
    default: return MG_ERR_ARG;
  }
  MG_HIP(hipGetLastError());
  ++g.launches;
  return MG_OK;
}

template <typename T>
int grp_laswp(Group& g, int64_t k1, int64_t k2, int64_t c0, int64_t c1, bool rhs) {
  typedef typename C2<T>::t T2;
  const int np = g.active(k1);
  if (np == 0 || c1 <= c0 || k2 <= k1) return MG_OK;
  GrpMap m;
  int total;
  if (!g.make_map(np, [&](int p) -> int64_t {
        const int64_t e = rhs ? c1 : std::min(c1, g.n[p]);
        return e > c0 ? (e - c0 + 255) / 256 : 0;
      }, m, total))
    return MG_OK;
  GrpSwp op{k1, k2, c0, c1, rhs ? 1 : 0};
  hipLaunchKernelGGL(k_grp_laswp<T2>, dim3(total), dim3(256), 0, g.s, g.dtab, m, op);
  MG_HIP(hipGetLastError());
  ++g.launches;
  return MG_OK;
}

template <typename T>
int grp_trsm_leaf(Group& g, bool upper, int64_t r0, int w, int64_t c0, int64_t c1, bool rhs) {
  const int np = g.active(r0);
  if (np == 0 || c1 <= c0) return MG_OK;
  GrpMap m;
  int total;
  if (!g.make_map(np, [&](int p) -> int64_t {
        const int64_t e = rhs ? c1 : std::min(c1, g.n[p]);
        return e > c0 ? (e - c0 + kTriWgCols - 1) / kTriWgCols : 0;
      }, m, total))
    return MG_OK;
  GrpTri op{r0, c0, c1, w, rhs ? 1 : 0};
  if (upper) hipLaunchKernelGGL((k_grp_trsm<T, true>), dim3(total), dim3(256), 0, g.s, g.dtab, m, op);
  else hipLaunchKernelGGL((k_grp_trsm<T, false>), dim3(total), dim3(256), 0, g.s, g.dtab, m, op);
  MG_HIP(hipGetLastError());
  ++g.launches;
  return MG_OK;
}

// C -= A B over the group; split: the complex-float trailing update on BF16x3
// products; tile: 0 = 128 x 64 (three-product FP64 / FP32), 1 = 64 x 64,
// 2 = 128 x 32 (three-product; products with few columns, as the residuals).
template <typename T>
int grp_gemm(Group& g, const GrpGemm& op, bool split, int tile) {
  const int np = g.active(op.k0);   // the depth needs rows beyond k0
  if (np == 0) return MG_OK;
  const int BM = tile == 1 ? 64 : 128;
  const int BN = split ? 128 : (tile == 2 ? 32 : 64);
  auto ext = [&](int64_t e, int64_t s, int64_t n, bool clamp) {
    const int64_t x = clamp ? std::min(e, n) : e;
    return x > s ? x - s : 0;
  };
  GrpMap m;
  int total;
  if (!g.make_map(np, [&](int p) -> int64_t {
        const int64_t M = ext(op.m1, op.m0, g.n[p], true), N = ext(op.n1, op.n0, g.n[p], !op.rhs),
                      K = ext(op.k1, op.k0, g.n[p], true);
        if (M == 0 || N == 0 || K == 0) return 0;
        return ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
      }, m, total))
    return MG_OK;
  if (split) {
    if (tile == 0) hipLaunchKernelGGL((k_grp_gemm_split<2, 128, 128, 4, 2>), dim3(total), dim3(512), 0, g.s, g.dtab, m, op);
    else hipLaunchKernelGGL((k_grp_gemm_split<2, 64, 128, 2, 2>), dim3(total), dim3(256), 0, g.s, g.dtab, m, op);
  } else if (tile == 0) {
    hipLaunchKernelGGL((k_grp_gemm<T, 128, 64, 16, 4, 2, true>), dim3(total), dim3(512), 0, g.s, g.dtab, m, op);
  } else if (tile == 2) {
    hipLaunchKernelGGL((k_grp_gemm<T, 128, 32, 16, 4, 2, true>), dim3(total), dim3(512), 0, g.s, g.dtab, m, op);
  } else {
    constexpr bool F64 = std::is_same<T, double>::value;
    hipLaunchKernelGGL((k_grp_gemm<T, 64, 64, 16, 2, 2, !F64>), dim3(total), dim3(256), 0, g.s, g.dtab, m, op);
  }
  MG_HIP(hipGetLastError());
  ++g.launches;
  return MG_OK;
}

// X := T^{-1} X for the diagonal block of order w at r0 (recursive above
// kGTri: halves, with the product between them).
template <typename T>
int grp_trsm(Group& g, bool upper, int64_t r0, int64_t w, int64_t c0, int64_t c1, bool rhs) {
  if (w <= kGTri) return grp_trsm_leaf<T>(g, upper, r0, (int)w, c0, c1, rhs);
  const int64_t w1 = split_half(w);
  if (!upper) {
    MG_TRY(grp_trsm<T>(g, false, r0, w1, c0, c1, rhs));
    MG_TRY(grp_gemm<T>(g, GrpGemm{r0 + w1, r0 + w, c0, c1, r0, r0 + w1, rhs ? 1 : 0}, false, 1));
    return grp_trsm<T>(g, false, r0 + w1, w - w1, c0, c1, rhs);
  }
  MG_TRY(grp_trsm<T>(g, true, r0 + w1, w - w1, c0, c1, rhs));
  MG_TRY(grp_gemm<T>(g, GrpGemm{r0, r0 + w1, c0, c1, r0 + w1, r0 + w, rhs ? 1 : 0}, false, 1));
  return grp_trsm<T>(g, true, r0, w1, c0, c1, rhs);
}

// Panel: columns [k, k + w) of every problem, rows [k, n).
template <typename T>
int grp_panel(Group& g, int64_t k, int64_t w) {
  if (w <= g.cfg.leaf) {
    if constexpr (std::is_same<T, float>::value)
      if (ts_leaf(g, k)) return grp_tslu(g, k, (int)w);
    return grp_getf2<T>(g, k, (int)w);
  }
  const int64_t w1 = split_half(w);
  MG_TRY(grp_panel<T>(g, k, w1));
  MG_TRY(grp_laswp<T>(g, k, k + w1, k + w1, k + w, false));
  MG_TRY(grp_trsm<T>(g, false, k, w1, k + w1, k + w, false));
  MG_TRY(grp_gemm<T>(g, GrpGemm{k + w1, INT64_MAX, k + w1, k + w, k, k + w1, 0}, false, w - w1 >= 64 ? 0 : 1));
  MG_TRY(grp_panel<T>(g, k + w1, w - w1));
  return grp_laswp<T>(g, k + w1, k + w, k, k + w1, false);
}

// Look-ahead of depth one: once step k has updated the next panel's
// columns, that panel is factored on the group's second stream while the
// rest of step k's trailing update runs on the call's stream (the two touch
// disjoint columns); step k+1 waits for the panel before its row swaps.
template <typename T>
int grp_getrf(Group& g, bool split) {
  const int64_t nmax = g.P ? g.n[0] : 0;
  hipStream_t s = g.s;
  hipLaunchKernelGGL(k_grp_zero_info, dim3((g.P + 255) / 256), dim3(256), 0, s, g.dtab, g.P);
  MG_HIP(hipGetLastError());
  const int64_t nb = g.cfg.nb;
  // BF16x3 with whole 128-column tiles: the updates read L21 and U12 split
  // once per step (the buffers are set up by the caller, outside any capture)
  const bool pre = split && dma_split(g);
  if (pre && !g.split) return MG_ERR_ARG;
  MG_TRY(grp_panel<T>(g, 0, nb));
  if (pre) MG_TRY(grp_pack(g, 0, true, 0));
  for (int64_t k = 0; k < nmax; k += nb) {
    const int64_t kn = k + nb;
    const int par = (int)((k / nb) & 1);
    if (kn < nmax) {
      MG_TRY(grp_laswp<T>(g, k, kn, kn, INT64_MAX, false));
      MG_TRY(grp_trsm<T>(g, false, k, nb, kn, INT64_MAX, false));
      if (pre) {
        MG_TRY(grp_pack(g, k, false, par));
        MG_TRY(grp_update_dma(g, k, 0, nb, par));
      } else {
        MG_TRY(grp_gemm<T>(g, GrpGemm{kn, INT64_MAX, kn, kn + nb, k, kn, 0}, split, g.cfg.tile));
      }
      MG_HIP(hipEventRecord(g.evA, s));
      MG_HIP(hipStreamWaitEvent(g.s2, g.evA, 0));
      g.s = g.s2;
      int st = grp_panel<T>(g, kn, nb);
      // the next step's L21, split while this step's update reads the other buffer
      if (st == MG_OK && pre) st = grp_pack(g, kn, true, par ^ 1);
      // this step's swaps on the columns left of it (disjoint from the
      // panel and from the rest of the update) ride on the panel's stream
      if (st == MG_OK && k > 0) st = grp_laswp<T>(g, k, kn, 0, k, false);
      g.s = s;
      MG_TRY(st);
      MG_HIP(hipEventRecord(g.evB, g.s2));
      if (pre) MG_TRY(grp_update_dma(g, k, nb, INT64_MAX, par));
      else MG_TRY(grp_gemm<T>(g, GrpGemm{kn, INT64_MAX, kn + nb, INT64_MAX, k, kn, 0}, split, g.cfg.tile));
      MG_HIP(hipStreamWaitEvent(s, g.evB, 0));
    } else if (k > 0) {
      MG_TRY(grp_laswp<T>(g, k, kn, 0, k, false));
    }
  }
  return MG_OK;
}

template <typename T>
int grp_getrs(Group& g) {
  const int64_t nmax = g.P ? g.n[0] : 0, nr = g.nrhs;
  if (nr == 0 || nmax == 0) return MG_OK;
  MG_TRY(grp_laswp<T>(g, 0, INT64_MAX, 0, nr, true));
  const int64_t sb = kGTri;
  const int tile = nr <= 32 ? 2 : 1;
  for (int64_t r0 = 0; r0 < nmax; r0 += sb) {
    MG_TRY(grp_trsm_leaf<T>(g, false, r0, (int)sb, 0, nr, true));
    MG_TRY(grp_gemm<T>(g, GrpGemm{r0 + sb, INT64_MAX, 0, nr, r0, r0 + sb, 1}, false, tile));
  }
  for (int64_t r0 = ((nmax - 1) / sb) * sb; r0 >= 0; r0 -= sb) {
    MG_TRY(grp_trsm_leaf<T>(g, true, r0, (int)sb, 0, nr, true));
    if (r0 > 0) MG_TRY(grp_gemm<T>(g, GrpGemm{0, r0, 0, nr, r0, r0 + sb, 1}, false, tile));
  }
  return MG_OK;
}

// BF16x3 groups with right-hand sides solve from inverted diagonal blocks.
bool inv_solve(const Group& g) { return g.prec == MG_BF16X3 && g.nrhs > 0 && g.solvebuf; }

int64_t sblocks(int64_t n) { return (n + kSB - 1) / kSB; }

// After the factorisation: perm and the inverted diagonal blocks.
int grp_solve_prep(Group& g) {
  hipLaunchKernelGGL(k_grp_perm, dim3(g.P), dim3(256), 0, g.s, g.dtab);
  MG_HIP(hipGetLastError());
  GrpMap m;
  int total;
  if (g.make_map(g.P, [&](int p) -> int64_t { return 2 * sblocks(g.n[p]); }, m, total)) {
    hipLaunchKernelGGL(k_grp_trinv, dim3(total), dim3(64), 0, g.s, g.dtab, m);
    MG_HIP(hipGetLastError());
  }
  g.launches += 2;
  return MG_OK;
}

int grp_solve_inv(Group& g) {
  const int nr = (int)g.nrhs, nct = (nr + 31) / 32;
  GrpMap m;
  int total;
  if (g.make_map(g.P, [&](int p) -> int64_t { return (g.n[p] * nr + 255) / 256; }, m, total)) {
    hipLaunchKernelGGL(k_grp_gather, dim3(total), dim3(256), 0, g.s, g.dtab, m, nr);
    MG_HIP(hipGetLastError());
    ++g.launches;
  }
  (void)nct;
  // counter set 0 cleared (it is clear after any completed solve: the
  // backward sweep clears it), the forward sweep on set 0 clears set 1 for
  // the backward sweep
  int* c0 = g.sweep_ctr;
  int* c1 = g.sweep_ctr + g.sweep_cap;
  hipLaunchKernelGGL(k_grp_clear, dim3(std::min(64, (g.sweep_cap + 255) / 256)), dim3(256), 0, g.s, c0, g.sweep_cap);
  hipLaunchKernelGGL(k_grp_sweep<true>, dim3(g.sweep_tickets), dim3(128), 0, g.s, g.dtab, g.sweep_lev, g.sweep_nlev,
                     g.sweep_coff, nr, c0, c1, g.sweep_cap);
  hipLaunchKernelGGL(k_grp_sweep<false>, dim3(g.sweep_tickets), dim3(128), 0, g.s, g.dtab, g.sweep_lev, g.sweep_nlev,
                     g.sweep_coff, nr, c1, c0, g.sweep_cap);
  MG_HIP(hipGetLastError());
  g.launches += 3;
  return MG_OK;
}

int run_factor(Group& g) {
  switch (g.prec) {
    case MG_C128: return grp_getrf<double>(g, false);
    case MG_C64: return grp_getrf<float>(g, false);
    case MG_BF16X3:
      MG_TRY(grp_getrf<float>(g, true));
      return inv_solve(g) ? grp_solve_prep(g) : MG_OK;
    default: return MG_ERR_ARG;
  }
}

int run_solve(Group& g) {
  if (g.prec == MG_C128) return grp_getrs<double>(g);
  return inv_solve(g) ? grp_solve_inv(g) : grp_getrs<float>(g);
}

// B -= A X for every problem (X: the auxiliary blocks), products in the
// group's element type.
int run_residual(Group& g) {
  if (g.prec == MG_C128) {   // the refinement's FP64 residual: one streaming pass over every matrix
    GrpMap m;
    int total;
    const int64_t nct = (g.nrhs + kResCols - 1) / kResCols;
    constexpr int RF = 2, KS = 2, WPE = 2;
    if (!g.make_map(g.P, [&](int p) -> int64_t { return (g.n[p] + 64 * RF - 1) / (64 * RF) * nct; }, m, total))
      return MG_OK;
    hipLaunchKernelGGL((k_grp_residual<RF, KS, WPE>), dim3(total), dim3(256), 0, g.s, g.dtab, m, (int)g.nrhs);
    MG_HIP(hipGetLastError());
    ++g.launches;
    return MG_OK;
  }
  const GrpGemm op{0, INT64_MAX, 0, g.nrhs, 0, INT64_MAX, 2};
  return grp_gemm<float>(g, op, false, g.nrhs <= 32 ? 2 : 0);
}

int run_colnorm(Group& g, double* out) {
  const dim3 grid((unsigned)((g.nrhs + 63) / 64), (unsigned)g.P);
  if (g.prec == MG_C128) hipLaunchKernelGGL(k_grp_colnorm<double2>, grid, dim3(256), 0, g.s, g.dtab, (int)g.nrhs, out);
  else hipLaunchKernelGGL(k_grp_colnorm<float2>, grid, dim3(256), 0, g.s, g.dtab, (int)g.nrhs, out);
  MG_HIP(hipGetLastError());
  ++g.launches;
  return MG_OK;
}

// ---- autotuning -------------------------------------------------------------
// The tunables are timed on a sample of the group (every k-th problem, at
// most kTuneProblems and kTuneBytes of copies, so the sample keeps the
// group's size mix) and the fastest is cached per (device, precision,
// ceil log2 of the largest order).
constexpr int kTuneProblems = 32;
constexpr int64_t kTuneBytes = 12ll << 30;

struct TuneKey {
  int dev, prec, nlog, pivot;
  bool operator<(const TuneKey& o) const {
    if (dev != o.dev) return dev < o.dev;
    if (prec != o.prec) return prec < o.prec;
    if (pivot != o.pivot) return pivot < o.pivot;
    return nlog < o.nlog;
  }
};

std::mutex g_tune_mu;
std::map<TuneKey, GrpCfg> g_tune_cache;

int ceil_log2(int64_t v) {
  int l = 0;
  while ((int64_t(1) << l) < v) ++l;
  return l;
}

TuneKey tune_key(const Group& g) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  return TuneKey{dev, g.prec, ceil_log2(g.P ? g.n[0] : 1), g.prec == MG_BF16X3 ? g.cfg.pivot : 0};
}

// panel widths around the measured best (256 for the 3D ensemble's orders
// 1500 .. 9000 in both precisions; the 1024-thread leaf and the 128 x 64
// tile won there at every width).  BF16x3 groups leave out 128: on a
// sample of 32 of the 256 problems it timed fastest, on the whole group
// slowest (1164 ms against 730 at 256 and 718 at 384).
const GrpCfg kCandidates[] = {{256, 16, 0, 1024}, {128, 16, 0, 1024}, {384, 16, 0, 1024}};
const GrpCfg kCandidatesSplit[] = {{256, 16, 0, 1024}, {384, 16, 0, 1024}};

int tune(Group& g, hipStream_t s, float* best_ms) {
  const size_t esz = g.prec == MG_C128 ? 16 : 8;
  // the sample: every stride-th problem (sorted order), within the byte cap
  int stride = (g.P + kTuneProblems - 1) / kTuneProblems;
  std::vector<int> pick;
  for (;;) {
    pick.clear();
    int64_t bytes = 0;
    for (int i = 0; i < g.P; i += stride) {
      pick.push_back(i);
      bytes += g.n[i] * g.n[i] * (int64_t)esz;
    }
    if (bytes <= kTuneBytes || stride >= g.P) break;
    stride *= 2;
  }
  Group t;
  t.prec = g.prec;
  t.P = (int)pick.size();
  t.s2 = g.s2;
  t.evA = g.evA;
  t.evB = g.evB;
  std::vector<void*> bufs;
  int st = MG_OK;
  int* ints = nullptr;
  int64_t nint = 0;
  for (int i : pick) nint += g.n[i] + 1;
  if (hipMalloc(&ints, nint * sizeof(int)) != hipSuccess) {
    set_error("group tuning: allocation failed");
    return MG_ERR_NOMEM;
  }
  int64_t ioff = 0;
  auto q_of = [&](int i) { return (int)(std::find(pick.begin(), pick.end(), i) - pick.begin()); };
  for (int i : pick) {
    void* a = nullptr;
    if (hipMalloc(&a, g.n[i] * g.n[i] * esz) != hipSuccess) {
      st = MG_ERR_NOMEM;
      set_error("group tuning: allocation failed");
      break;
    }
    bufs.push_back(a);
    GrpProb e{a, g.n[i], g.n[i], ints + ioff, ints + ioff + g.n[i], nullptr, 0, nullptr, 0, (int64_t)q_of(i)};
    ioff += g.n[i] + 1;
    t.host.push_back(e);
    t.n.push_back(g.n[i]);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (st == MG_OK && (hipMalloc(&t.dtab, sizeof(GrpProb) * t.P) != hipSuccess ||
                      hipMemcpy(t.dtab, t.host.data(), sizeof(GrpProb) * t.P, hipMemcpyHostToDevice) != hipSuccess ||
                      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
    st = MG_ERR_HIP;
    set_error("group tuning: set-up failed");
  }
  GrpCfg best = g.cfg;
  float bms = -1.f;
  auto refill = [&]() -> int {
    for (size_t q = 0; q < pick.size(); ++q) {
      const GrpProb& src = g.host[pick[q]];
      MG_HIP(hipMemcpy2DAsync(bufs[q], g.n[pick[q]] * esz, src.A, src.lda * esz, g.n[pick[q]] * esz,
                              g.n[pick[q]], hipMemcpyDeviceToDevice, s));
    }
    return MG_OK;
  };
  t.s = s;
  const bool sp = g.prec == MG_BF16X3;
  const GrpCfg* cands = sp ? kCandidatesSplit : kCandidates;
  const int nc = sp ? (int)(sizeof(kCandidatesSplit) / sizeof(kCandidatesSplit[0]))
                    : (int)(sizeof(kCandidates) / sizeof(kCandidates[0]));
  for (int c = -1; c < nc && st == MG_OK; ++c) {   // c = -1: untimed warm-up
    t.cfg = cands[c < 0 ? 0 : c];
    t.cfg.pivot = g.cfg.pivot;
    if ((st = ensure_split(t)) != MG_OK) break;
    if ((st = refill()) != MG_OK) break;
    if (hipEventRecord(e0, s) != hipSuccess) { st = MG_ERR_HIP; break; }
    if ((st = run_factor(t)) != MG_OK) break;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { st = MG_ERR_HIP; break; }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (c >= 0 && (bms < 0 || ms < bms)) {
      bms = ms;
      best = t.cfg;
    }
  }
  (void)hipStreamSynchronize(s);
  (void)hipStreamSynchronize(g.s2);
  for (void* b : bufs) (void)hipFree(b);
  if (t.split) (void)hipFree(t.split);
  if (t.cand) (void)hipFree(t.cand);
  if (t.solvebuf) (void)hipFree(t.solvebuf);
  (void)hipFree(ints);
  if (t.dtab) (void)hipFree(t.dtab);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st != MG_OK) return st;
  g.cfg = best;
  if (best_ms) *best_ms = bms;
  return MG_OK;
}

// Record op (0: factor, 1: solve) into the group's graph slot: the launch
// sequence is captured on the group's capture stream (nothing runs) and
// instantiated; replay launches it on any stream.  The sequence holds no
// memset nodes (info is cleared by a kernel).
int capture(Group& g, int op) {
  if (g.exec[op]) {
    (void)hipGraphExecDestroy(g.exec[op]);
    g.exec[op] = nullptr;
  }
  if (g.graph[op]) {
    (void)hipGraphDestroy(g.graph[op]);
    g.graph[op] = nullptr;
  }
  if (op == 0) MG_TRY(ensure_split(g));   // allocates: not inside the capture
  MG_HIP(hipStreamBeginCapture(g.scap, hipStreamCaptureModeThreadLocal));
  g.s = g.scap;
  const int st = op == 0 ? run_factor(g) : run_solve(g);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(g.scap, &graph);
  if (st != MG_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return st;
  }
  if (e != hipSuccess) {
    set_error("group capture: %s", hipGetErrorString(e));
    return MG_ERR_HIP;
  }
  g.graph[op] = graph;
  MG_HIP(hipGraphInstantiate(&g.exec[op], graph, nullptr, nullptr, 0));
  return MG_OK;
}

}  // namespace
}  // namespace mg

struct mg_group {
  mg::Group g;
};

extern "C" {

int mg_group_destroy(mg_group* h);

int mg_group_create(int prec, int count, void* const* A, const int64_t* n, const int64_t* lda, int* const* ipiv,
                    int* const* info, void* const* B, const int64_t* ldb, int64_t nrhs, mg_group** out) {
  if (!out || count < 0 || count > mg::kMaxGroup || nrhs < 0) return MG_ERR_ARG;
  if (prec != MG_C128 && prec != MG_C64 && prec != MG_BF16X3) return MG_ERR_ARG;
  if (count > 0 && (!A || !n || !lda || !ipiv || !info)) return MG_ERR_ARG;
  if (nrhs > 0 && count > 0 && (!B || !ldb)) return MG_ERR_ARG;
  for (int i = 0; i < count; ++i) {
    if (!A[i] || !ipiv[i] || !info[i] || n[i] < 1 || lda[i] < n[i] || n[i] > INT_MAX / 2) return MG_ERR_ARG;
    if (nrhs > 0 && (!B[i] || ldb[i] < nrhs)) return MG_ERR_ARG;
  }
  mg_group* h = new (std::nothrow) mg_group();
  if (!h) return MG_ERR_NOMEM;
  mg::Group& g = h->g;
  g.prec = prec;
  g.P = count;
  g.nrhs = nrhs;
  g.order.resize(count);
  std::iota(g.order.begin(), g.order.end(), 0);
  std::stable_sort(g.order.begin(), g.order.end(), [&](int a, int b) { return n[a] > n[b]; });
  for (int i = 0; i < count; ++i) {
    const int c = g.order[i];
    mg::GrpProb e{A[c], lda[c], n[c], ipiv[c], info[c], nrhs > 0 ? B[c] : nullptr, nrhs > 0 ? ldb[c] : 0,
                  nullptr, 0, c};
    g.host.push_back(e);
    g.n.push_back(n[c]);
  }
  if (hipStreamCreateWithFlags(&g.s2, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&g.scap, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&g.evA, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&g.evB, hipEventDisableTiming) != hipSuccess) {
    mg_group_destroy(h);
    mg::set_error("mg_group_create: stream or event creation failed");
    return MG_ERR_HIP;
  }
  if (count > 0) {
    if (hipMalloc(&g.dtab, sizeof(mg::GrpProb) * count) != hipSuccess) {
      g.dtab = nullptr;
      mg_group_destroy(h);
      mg::set_error("mg_group_create: table allocation failed");
      return MG_ERR_NOMEM;
    }
    if (hipMemcpy(g.dtab, g.host.data(), sizeof(mg::GrpProb) * count, hipMemcpyHostToDevice) != hipSuccess) {
      mg_group_destroy(h);
      mg::set_error("mg_group_create: table upload failed");
      return MG_ERR_HIP;
    }
  }
  *out = h;
  return MG_OK;
}

int mg_group_destroy(mg_group* h) {
  if (!h) return MG_OK;
  if (h->g.s2) (void)hipStreamSynchronize(h->g.s2);
  for (int op = 0; op < 2; ++op) {
    if (h->g.exec[op]) (void)hipGraphExecDestroy(h->g.exec[op]);
    if (h->g.graph[op]) (void)hipGraphDestroy(h->g.graph[op]);
  }
  if (h->g.scap) (void)hipStreamDestroy(h->g.scap);
  if (h->g.dtab) (void)hipFree(h->g.dtab);
  if (h->g.split) (void)hipFree(h->g.split);
  if (h->g.cand) (void)hipFree(h->g.cand);
  if (h->g.solvebuf) (void)hipFree(h->g.solvebuf);
  if (h->g.evA) (void)hipEventDestroy(h->g.evA);
  if (h->g.evB) (void)hipEventDestroy(h->g.evB);
  if (h->g.s2) (void)hipStreamDestroy(h->g.s2);
  delete h;
  return MG_OK;
}

int mg_group_configure(mg_group* h, int nb, int leaf, int tile, int leaf_threads, int pivot) {
  if (!h) return MG_ERR_ARG;
  if (nb < 16 || nb > 1024 || nb % 16 != 0 || leaf < 1 || leaf > kGLeaf || (tile != 0 && tile != 1) ||
      (leaf_threads != 256 && leaf_threads != 512 && leaf_threads != 1024) || (pivot != 0 && pivot != 1))
    return MG_ERR_ARG;
  h->g.cfg.nb = nb;
  h->g.cfg.leaf = leaf;
  h->g.cfg.tile = tile;
  h->g.cfg.lthreads = leaf_threads;
  h->g.cfg.pivot = pivot;
  return MG_OK;
}

int mg_group_getrf(mg_group* h, void* stream) {
  if (!h) return MG_ERR_ARG;
  MG_TRY(mg::ensure_split(h->g));
  h->g.s = reinterpret_cast<hipStream_t>(stream);
  return mg::run_factor(h->g);
}

int mg_group_getrs(mg_group* h, void* stream) {
  if (!h) return MG_ERR_ARG;
  h->g.s = reinterpret_cast<hipStream_t>(stream);
  return mg::run_solve(h->g);
}

int64_t mg_group_launches(const mg_group* h) { return h ? h->g.launches : -1; }

int mg_group_tune(mg_group* h, int* cfg_out, float* best_ms, void* stream) {
  if (!h) return MG_ERR_ARG;
  mg::Group& g = h->g;
  const mg::TuneKey key = mg::tune_key(g);
  int st = MG_OK;
  {
    std::lock_guard<std::mutex> lock(mg::g_tune_mu);
    auto it = mg::g_tune_cache.find(key);
    if (it != mg::g_tune_cache.end()) {
      g.cfg = it->second;
      if (best_ms) *best_ms = 0.f;
    } else {
      st = mg::tune(g, reinterpret_cast<hipStream_t>(stream), best_ms);
      if (st == MG_OK) mg::g_tune_cache[key] = g.cfg;
    }
  }
  if (st == MG_OK && cfg_out) {
    cfg_out[0] = g.cfg.nb;
    cfg_out[1] = g.cfg.leaf;
    cfg_out[2] = g.cfg.tile;
    cfg_out[3] = g.cfg.lthreads;
  }
  return st;
}

int mg_group_set_aux(mg_group* h, void* const* X, const int64_t* ldx) {
  if (!h || !X || !ldx) return MG_ERR_ARG;
  mg::Group& g = h->g;
  if (g.nrhs == 0) return MG_ERR_ARG;
  for (int i = 0; i < g.P; ++i) {
    const int c = g.order[i];
    if (!X[c] || ldx[c] < g.nrhs) return MG_ERR_ARG;
    g.host[i].X = X[c];
    g.host[i].ldx = ldx[c];
  }
  if (g.P > 0) MG_HIP(hipMemcpy(g.dtab, g.host.data(), sizeof(mg::GrpProb) * g.P, hipMemcpyHostToDevice));
  return MG_OK;
}

int mg_group_residual(mg_group* h, void* stream) {
  if (!h || h->g.nrhs == 0) return MG_ERR_ARG;
  for (const mg::GrpProb& e : h->g.host)
    if (!e.X) {
      mg::set_error("mg_group_residual: no auxiliary blocks (mg_group_set_aux)");
      return MG_ERR_ARG;
    }
  h->g.s = reinterpret_cast<hipStream_t>(stream);
  return mg::run_residual(h->g);
}

int mg_group_colnorm(mg_group* h, double* norms, void* stream) {
  if (!h || !norms || h->g.nrhs == 0) return MG_ERR_ARG;
  h->g.s = reinterpret_cast<hipStream_t>(stream);
  return mg::run_colnorm(h->g, norms);
}

int mg_group_capture(mg_group* h, int op) {
  if (!h || (op != 0 && op != 1)) return MG_ERR_ARG;
  if (op == 1 && h->g.nrhs == 0) return MG_ERR_ARG;
  return mg::capture(h->g, op);
}

int mg_group_replay(mg_group* h, int op, void* stream) {
  if (!h || (op != 0 && op != 1)) return MG_ERR_ARG;
  if (!h->g.exec[op]) {
    mg::set_error("mg_group_replay: nothing captured for op %d", op);
    return MG_ERR_ARG;
  }
  MG_HIP(hipGraphLaunch(h->g.exec[op], reinterpret_cast<hipStream_t>(stream)));
  return MG_OK;
}

}  // extern "C"
