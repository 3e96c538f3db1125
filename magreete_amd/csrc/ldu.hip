// Block LDU factorisation of a complex-float matrix without pivoting: the
// preconditioner of the mixed-precision solve.
//
// With 256 x 256 diagonal blocks, A = Lh D Uh, Lh (Uh) block unit lower
// (upper) triangular and D = blockdiag(S_k), S_k the k-th diagonal block of
// the successive Schur complements.  Step k of the factorisation:
//   S_k^{-1}            eight workgroups (one 32-column slab each, held in
//                       registers), in-place Gauss-Jordan with 32-column
//                       panels (BF16x3 MFMA rank-32 updates, FP32 pivots),
//                       plus one FP32 Newton-Schulz step when the other
//                       products are FP32;
//   Uh_k = S_k^{-1} A12, Lh_k = A21 S_k^{-1}     stored over A12 and A21;
//   A22 -= A21 S_k^{-1} A12                      the trailing update.
// A solve is then GEMMs only: the forward sweep with Lh, one batched
// product with the S_k^{-1}, the backward sweep with Uh; no triangular
// kernels and no per-column dependency chains.  Each sweep is either one
// GEMM launch per block column or, with split-bf16 products, one dataflow
// launch (k_ldu_sweep).
// Look-ahead: once step k has updated block row and column k+1, step k+1's
// inverse and panels run on the workspace's second stream while the rest of
// step k's trailing update runs on the caller's stream.
// Nothing here pivots: a zero pivot inside a diagonal block is reported in
// info (first one, 1-based row), and the refinement that uses these factors
// falls back to a pivoted factorisation when they are not good enough.
#include "bf16x3.h"
#include "cgemm.h"
#include "dma_update.h"
#include "lu.h"
#include "mg_internal.h"
#include "psplit.h"
#include "split_gemm.h"

namespace {

constexpr int kIB = 256;          // diagonal block order (= kInvBlock)
constexpr int kPW = 32;           // Gauss-Jordan panel width
constexpr int kPK = kPW + 8;      // padded k stride of the bf16 planes
static_assert(kIB == 8 * kPW, "slab layout of k_block_inverse");

__device__ inline float2 cmulf(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline float2 crecipf(float2 a) {
  const float d = a.x * a.x + a.y * a.y;
  return make_float2(a.x / d, -a.y / d);
}

__device__ inline void split4(float2 v, unsigned short& rh, unsigned short& rl, unsigned short& ih,
                              unsigned short& il) {
  mg_split(v.x, rh, rl);
  mg_split(v.y, ih, il);
}

constexpr int kSlabs = kIB / kPW;     // workgroups of the block inverse: one per 32-column slab
constexpr int kMwThreads = 256;       // 4 waves; wave w holds slab rows 64w .. 64w+63
constexpr long long kInvSpin = 1LL << 22;

struct alignas(16) InvSmem {
  unsigned short pa[4][kPW][kPK];     // P^{-1} as MFMA A operand: [plane][row][k]
  unsigned short xb[4][kPW][kPK];     // the slab's pivot rows as B operand: [plane][column][k]
  unsigned short rb[4][kPW][kPK];     // R (the slab's new pivot rows) as B operand
  float2 pf[kPW][kPW + 1];            // P^{-1} (and the even Gauss-Jordan steps' input)
  float2 pg[kPW][kPW + 1];            // the odd steps' input
  float2 rf[kPW][kPW + 1];            // R
};

__device__ inline void split_planes(float2 v, unsigned short* p0, unsigned short* p1, unsigned short* p2,
                                    unsigned short* p3) {
  split4(v, *p0, *p1, *p2, *p3);
}

// Barrier over the kSlabs workgroups of k_block_inverse (counter bar, target
// the arrivals so far).  Every wave's stores complete before the arrival; a
// wait past the spin bound (or another workgroup's) sets err and returns.
__device__ inline void inv_barrier(unsigned* bar, unsigned target, int* err) {
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    long long spins = 0;
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      if (++spins > kInvSpin) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// X (256 x 256, row stride 256) <- inverse of the h x h block A padded with
// the identity, by in-place Gauss-Jordan without pivoting over 32-column
// panels.  kSlabs workgroups; workgroup q keeps columns 32q .. 32q+31 of X
// in its MFMA accumulators for the whole inversion.  Panel p:
//   slab p, published by its owner to exch (coherent stores), gives the
//   pivot block P (rows 32p..) and the column panel C;
//   every workgroup inverts P in LDS (FP32 Gauss-Jordan, all threads);
//   R = P^{-1} X_p (the slab's pivot rows; R = P^{-1} on slab p) becomes the
//   slab's pivot rows, and every other row gets X_i <- X_i - C_i R (slab p:
//   X_i <- -C_i P^{-1}) as BF16x3 MFMA products.
// One barrier per panel; zero pivots are reported in info (first, 1-based).
// bar / err: this launch's barrier counter and failure flag (zero on entry);
// next: the other set, cleared here for the next launch.  One wave per SIMD
// (its eight workgroups take eight CUs): the register budget of two spilled
// 80 bytes per lane; with one, no spill and the 12288-row BF16x3
// factorisation 20.4 -> 20.0 ms.
__global__ __launch_bounds__(kMwThreads) __attribute__((amdgpu_waves_per_eu(1))) void k_block_inverse(const float2* __restrict__ A, int64_t lda, int h,
                                                              float2* __restrict__ X, int* info, int64_t row0,
                                                              float2* exch, unsigned* bar, int* err, int* next) {
  __shared__ InvSmem sm;
  const int q = blockIdx.x;
  if (q == 0 && threadIdx.x == 0) {   // the next launch's barrier counter and flag
    __hip_atomic_store(next, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(next + 32, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  // slab item r of tile (a, b): row 64 wave + 16 a + 4 lg + r, column 32 q + 16 b + li
  v4f xr[4][2], xi[4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 64 * wave + 16 * a + 4 * lg + r, col = kPW * q + 16 * b + li;
        float2 v = make_float2(row == col ? 1.f : 0.f, 0.f);
        if (row < h && col < h) v = A[(int64_t)row * lda + col];
        xr[a][b][r] = v.x;
        xi[a][b][r] = v.y;
      }
  auto publish = [&](float2* buf) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          mg_sg::coherent_store(&buf[(64 * wave + 16 * a + 4 * lg + r) * kPW + 16 * b + li],
                                make_float2(xr[a][b][r], xi[a][b][r]));
  };
  unsigned arrivals = 0;
  if (q == 0) publish(exch);
  inv_barrier(bar, (++arrivals) * kSlabs, err);

  for (int p = 0; p < kSlabs; ++p) {
    const float2* C = exch + (p & 1) * (kIB * kPW);   // slab p: 256 x 32, row stride 32
    // (1) P^{-1} by Gauss-Jordan in LDS, all threads: thread (row pi, columns
    // pj .. pj + 3); step st reads one buffer and writes the other (pf on
    // even steps' input, pg on odd ones'; 32 steps end in pf), one barrier
    // per step (two with one buffer; the 12288-row BF16x3 factorisation
    // 20.5 -> 20.3 ms)
    {
      const int pi = tid >> 3, pj = (tid & 7) * 4;
#pragma unroll
      for (int e = 0; e < 4; ++e) sm.pf[pi][pj + e] = mg_sg::coherent_load(&C[(kPW * p + pi) * kPW + pj + e]);
      __syncthreads();
      static_assert(kPW % 2 == 0, "the last step writes pf");
      for (int st = 0; st < kPW; ++st) {
        float2(*in)[kPW + 1] = (st & 1) ? sm.pg : sm.pf;
        float2(*out)[kPW + 1] = (st & 1) ? sm.pf : sm.pg;
        const float2 piv = in[st][st], ci = in[pi][st];
        float2 rj[4], xv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          rj[e] = in[st][pj + e];
          xv[e] = in[pi][pj + e];
        }
        if (q == 0 && tid == 0 && piv.x == 0.f && piv.y == 0.f && kPW * p + st < h)
          atomicCAS(info, 0, (int)(row0 + kPW * p + st + 1));
        const float rd = __builtin_amdgcn_rcpf(piv.x * piv.x + piv.y * piv.y);
        const float2 inv = make_float2(piv.x * rd, -piv.y * rd);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int j = pj + e;
          const float2 u = (j == st) ? inv : cmulf(rj[e], inv);   // row st of the result
          const float2 tq = cmulf(ci, (j == st) ? inv : u);
          out[pi][j] = (pi == st) ? u : (j == st) ? make_float2(-tq.x, -tq.y)
                                                 : make_float2(xv[e].x - tq.x, xv[e].y - tq.y);
        }
        __syncthreads();
      }
#pragma unroll
      for (int e = 0; e < 4; ++e)
        split_planes(sm.pf[pi][pj + e], &sm.pa[0][pi][pj + e], &sm.pa[1][pi][pj + e], &sm.pa[2][pi][pj + e],
                     &sm.pa[3][pi][pj + e]);
    }
    // the slab's pivot rows (wave pw, row tiles pa0 and pa0 + 1) as B operand
    const int pw = p >> 1, pa0 = 2 * (p & 1);
    if (q != p && wave == pw) {
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int k = 16 * t + 4 * lg + r, c = 16 * b + li;
            const int a = pa0 + t;
            split_planes(make_float2(xr[a][b][r], xi[a][b][r]), &sm.xb[0][c][k], &sm.xb[1][c][k], &sm.xb[2][c][k],
                         &sm.xb[3][c][k]);
          }
    }
    __syncthreads();
    // (2) this wave's rows of C (the A operand of the update), loaded after the
    // barrier so that their registers are not live across the pivot inversion
    float2 cf[4][8];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int e = 0; e < 8; ++e) cf[a][e] = mg_sg::coherent_load(&C[(64 * wave + 16 * a + li) * kPW + 8 * lg + e]);
    // (3) R = P^{-1} X_p (q != p) or P^{-1} (q == p); wave w computes tile (w / 2, w % 2)
    {
      const int ti = wave >> 1, tj = wave & 1;
      float2 rv[4];
      if (q == p) {
#pragma unroll
        for (int r = 0; r < 4; ++r) rv[r] = sm.pf[16 * ti + 4 * lg + r][16 * tj + li];
      } else {
        const int ar = 16 * ti + li, bc = 16 * tj + li;
        v4f r = {0.f, 0.f, 0.f, 0.f}, im = {0.f, 0.f, 0.f, 0.f};
        mg_cmac12(mg_ld8(&sm.pa[0][ar][lg * 8]), mg_ld8(&sm.pa[1][ar][lg * 8]), mg_ld8(&sm.pa[2][ar][lg * 8]),
                  mg_ld8(&sm.pa[3][ar][lg * 8]), mg_ld8(&sm.xb[0][bc][lg * 8]), mg_ld8(&sm.xb[1][bc][lg * 8]),
                  mg_ld8(&sm.xb[2][bc][lg * 8]), mg_ld8(&sm.xb[3][bc][lg * 8]), r, im);
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) rv[rr] = make_float2(r[rr], im[rr]);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = 16 * ti + 4 * lg + r, c = 16 * tj + li;
        sm.rf[k][c] = rv[r];
        split_planes(rv[r], &sm.rb[0][c][k], &sm.rb[1][c][k], &sm.rb[2][c][k], &sm.rb[3][c][k]);
      }
    }
    __syncthreads();
    // (4) slab update: pivot rows <- R; others X_i <- X_i - C_i R (slab p: -C_i P^{-1})
    bf16x8 bp[2][4];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) bp[b][pl] = mg_ld8(&sm.rb[pl][16 * b + li][lg * 8]);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      if (wave == pw && (a == pa0 || a == pa0 + 1)) {
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float2 v = sm.rf[16 * (a - pa0) + 4 * lg + r][16 * b + li];
            xr[a][b][r] = v.x;
            xi[a][b][r] = v.y;
          }
        continue;
      }
      unsigned short ph[4][8];   // -C rows of tile a as bf16x3 planes
#pragma unroll
      for (int e = 0; e < 8; ++e) split4(make_float2(-cf[a][e].x, -cf[a][e].y), ph[0][e], ph[1][e], ph[2][e], ph[3][e]);
      bf16x8 na[4];
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) na[pl] = mg_as_bf16x8(mg_pack8(ph[pl]));
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        if (q == p) {
          xr[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
          xi[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
        }
        mg_cmac12(na[0], na[1], na[2], na[3], bp[b][0], bp[b][1], bp[b][2], bp[b][3], xr[a][b], xi[a][b]);
      }
    }
    // (5) the next panel's slab, published by its owner
    if (p + 1 < kSlabs) {
      if (q == p + 1) publish(exch + ((p + 1) & 1) * (kIB * kPW));
      inv_barrier(bar, (++arrivals) * kSlabs, err);
    } else {
      __syncthreads();
    }
  }
  const bool bad = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
  const float nan = __builtin_nanf("");
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 64 * wave + 16 * a + 4 * lg + r, col = kPW * q + 16 * b + li;
        X[row * kIB + col] = bad ? make_float2(nan, nan) : make_float2(xr[a][b][r], xi[a][b][r]);
      }
}

__global__ void k_zero_info(int* p) { *p = 0; }

__global__ void k_zero_ints(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Split precisions (BF16x3, BF16x6): the factor's panels are stored pre-split
// (psplit.h), once per factorisation, after the D inverses.  With the matrix
// padded to np = nblk 256 rows, step k < nblk - 1 has rows(k) = np - (k+1) 256
// and two stored operands of rows(k) 256 elements each:
//   SPL(k):  Lh's block column k (rows (k+1) 256 .. np), A-operand units,
//            32 groups per row (the forward sweep's and the trailing update's
//            A operand);
//   SPUr(k): Uh's block row k (256 rows; k = columns (k+1) 256 .. np),
//            A-operand units, rows(k) / 8 groups per row (the backward
//            sweep's A operand).
// Each is 32 U rows(k) units at unit offset 32 U sum_{k' < k} rows(k')
// (U = 2 NS units per row and group); all of SPL comes first, then SPUr.
struct SplitLayout {
  int64_t np, nblk;
  int U;
  __host__ __device__ int64_t rows(int64_t k) const { return np - (k + 1) * kIB; }
  __host__ __device__ int64_t off(int64_t k) const {
    return (int64_t)(kIB / 8) * U * (k * np - (int64_t)kIB * k * (k + 1) / 2);
  }
  __host__ __device__ int64_t total() const { return off(nblk); }
};

// A-operand units of rows [0, rows_pad) x groups [0, ldg) from A (rows x K,
// row-major lda); zeros outside rows x K.
template <int NS>
__global__ void k_pack_rows(const float2* __restrict__ A, int64_t lda, int64_t rows, int64_t K, int64_t rows_pad,
                            int64_t ldg, u32x4* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows_pad * ldg) return;
  const int64_t r = t / ldg, g = t - r * ldg;
  float2 v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int64_t k = g * 8 + e;
    v[e] = (r < rows && k < K) ? A[r * lda + k] : make_float2(0.f, 0.f);
  }
  u32x4 u[2 * NS];
  mg_ps::split8<NS>(v, u);
#pragma unroll
  for (int q = 0; q < 2 * NS; ++q) out[t * 2 * NS + q] = u[q];
}

// B-operand units of columns [0, cols_pad) x groups [0, ldg) from B (K x
// cols, row-major ldb); zeros outside K x cols.  Consecutive threads take
// consecutive columns (coalesced reads).
template <int NS>
__global__ void k_pack_cols(const float2* __restrict__ B, int64_t ldb, int64_t K, int64_t cols, int64_t cols_pad,
                            int64_t ldg, u32x4* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cols_pad * ldg) return;
  const int64_t g = t / cols_pad, c = t - g * cols_pad;
  float2 v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int64_t k = g * 8 + e;
    v[e] = (c < cols && k < K) ? B[k * ldb + c] : make_float2(0.f, 0.f);
  }
  u32x4 u[2 * NS];
  mg_ps::split8<NS>(v, u);
#pragma unroll
  for (int q = 0; q < 2 * NS; ++q) out[(c * ldg + g) * 2 * NS + q] = u[q];
}

template <int NS>
int pack_rows(const float2* A, int64_t lda, int64_t rows, int64_t K, int64_t rows_pad, int64_t ldg, u32x4* out,
              hipStream_t s) {
  const int64_t t = rows_pad * ldg;
  if (t <= 0) return MG_OK;
  hipLaunchKernelGGL((k_pack_rows<NS>), dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s, A, lda, rows, K, rows_pad,
                     ldg, out);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

template <int NS>
int pack_cols(const float2* B, int64_t ldb, int64_t K, int64_t cols, int64_t cols_pad, int64_t ldg, u32x4* out,
              hipStream_t s) {
  const int64_t t = cols_pad * ldg;
  if (t <= 0) return MG_OK;
  hipLaunchKernelGGL((k_pack_cols<NS>), dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s, B, ldb, K, cols, cols_pad,
                     ldg, out);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

// A panel of the split factorisation, C = A B (complex float operands, row-major;
// entries outside M x N x K read as 0), written pre-split as A-operand units
// (rows of C, k along its columns, ldg groups per row) over the padded
// extent Mp x Np (Np a multiple of 8) of the grid, whose tiles past M x N hold
// zeros.  The accumulators go through LDS so that every thread writes whole
// 8-column groups.
template <int NS>
__global__ __launch_bounds__(256) void k_panel_split(int64_t M, int64_t N, int64_t K, const float2* __restrict__ A,
                                                     int64_t lda, const float2* __restrict__ B, int64_t ldb,
                                                     int64_t Mp, int64_t Np, u32x4* __restrict__ out, int64_t ldg,
                                                     int nbn) {
  constexpr int BM = 64, BN = 64;
  typedef mg_sg::Tile<NS, BM, BN, 2, 2> Tile;
  __shared__ typename Tile::S sm;
  static_assert(sizeof(typename Tile::S) >= sizeof(float2) * BM * (BN + 1), "staging fits the operand LDS");
  const int64_t i0 = (int64_t)(blockIdx.x / nbn) * BM, j0 = (int64_t)(blockIdx.x % nbn) * BN;
  Tile t;
  t.init(nullptr, 0, 0, 0, make_float2(0.f, 0.f), false);
  t.mac(sm, A + i0 * lda, lda, B + j0, ldb, M - i0, N - j0, K);
  __syncthreads();
  float2* stage = reinterpret_cast<float2*>(&sm);   // [BM][BN + 1]
#pragma unroll
  for (int a = 0; a < Tile::TM; ++a)
#pragma unroll
    for (int b = 0; b < Tile::TN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        stage[Tile::row_of(a, r) * (BN + 1) + Tile::col_of(b)] = make_float2(t.cr[a][b][r], t.ci[a][b][r]);
  __syncthreads();
  for (int c = threadIdx.x; c < BM * (BN / 8); c += 256) {
    const int row = c / (BN / 8), g = c % (BN / 8);
    const int64_t grow = i0 + row, gcol = j0 + 8 * g;
    if (grow >= Mp || gcol >= Np) continue;
    float2 v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = stage[row * (BN + 1) + 8 * g + e];
    u32x4 u[2 * NS];
    mg_ps::split8<NS>(v, u);
    u32x4* dst = out + (grow * ldg + gcol / 8) * 2 * NS;
#pragma unroll
    for (int q = 0; q < 2 * NS; ++q) dst[q] = u[q];
  }
}

template <int NS>
int panel_split(int64_t M, int64_t N, int64_t K, const float2* A, int64_t lda, const float2* B, int64_t ldb,
                int64_t Mp, int64_t Np, u32x4* out, int64_t ldg, hipStream_t s) {
  if (Mp <= 0 || Np <= 0) return MG_OK;
  const int nbm = (int)((Mp + 63) / 64), nbn = (int)((Np + 63) / 64);
  hipLaunchKernelGGL(k_panel_split<NS>, dim3(nbm * nbn), dim3(256), 0, s, M, N, K, A, lda, B, ldb, Mp, Np, out, ldg,
                     nbn);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

// Up to four (A, B) operand pairs of the split trailing update.
constexpr int kMaxSrc = 4;
struct UpdSrc {
  const u32x4* A[kMaxSrc];
  const u32x4* B[kMaxSrc];
  int n;
};

// Trailing update of the split factorisation, C -= sum_s A_s B_s with every A
// (M rows) and B (N columns) pre-split over one 256-block of k (the panels of
// one to four steps); C complex float row-major.  The accumulators start from
// -C (loaded before the k-loop).  The pre-split panels hold whole 256-row
// blocks (zeros past n) and every update starts on a block boundary, so the
// operand tiles are read without bounds checks.
template <int NS, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void k_ldu_update(int64_t M, int64_t N, UpdSrc src,
                                                             float2* __restrict__ C, int64_t ldc, int nbm, int nbn) {
  typedef mg_ps::PTile<NS, BM, BN, WM, WN> Tile;
  constexpr int U = 2 * NS, KG = kIB / 8;
  __shared__ typename Tile::S sm;
  int tm, tn;
  mg_tile_order(blockIdx.x, nbm, nbn, tm, tn);
  const int64_t i0 = (int64_t)tm * BM, j0 = (int64_t)tn * BN;
  const u32x4* A[kMaxSrc];
  const u32x4* B[kMaxSrc];
#pragma unroll
  for (int q = 0; q < kMaxSrc; ++q) {
    A[q] = src.A[q] + i0 * KG * U;
    B[q] = src.B[q] + j0 * KG * U;
  }
  Tile t;
  t.init(C + i0 * ldc + j0, ldc, M - i0, N - j0, make_float2(-1.f, 0.f), true);
  static_assert(kIB % BM == 0 && kIB % BN == 0, "tiles within the padded 256-row blocks");
  t.template pmac_n<true>(sm, A, B, src.n, KG, M - i0, N - j0, KG);
  t.store(C + i0 * ldc + j0, ldc, M - i0, N - j0, make_float2(-1.f, 0.f), make_float2(1.f, 0.f), false);
}

template <int NS, int BM, int BN, int WM, int WN>
int launch_update(int64_t M, int64_t N, const UpdSrc& src, float2* C, int64_t ldc, hipStream_t s) {
  const int nbm = (int)((M + BM - 1) / BM), nbn = (int)((N + BN - 1) / BN);
  hipLaunchKernelGGL((k_ldu_update<NS, BM, BN, WM, WN>), dim3(nbm * nbn), dim3(64 * WM * WN), 0, s, M, N, src, C, ldc,
                     nbm, nbn);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

// ---- BF16x3 trailing update for large C: operand tiles copied to LDS by DMA --
// 128 x 128 tiles (dma_update.h: the same per-wave MFMA work and accumulation
// order as k_ldu_update; C enters the sum first, as there, or last, in the
// LATE form), the k-loop over NSRC steps of 8 k-tiles.  Standalone,
// 12032^2 x 256 with C first: 1.03 ms against 1.19 for k_ldu_update; x 512:
// 1.67 against 1.70 (the same bits).
namespace dma {
using mg_dma::BM;
using mg_dma::BN;
using mg_dma::U;

template <int NSRC, bool LATE>
__global__ __launch_bounds__(mg_dma::NT) void k_ldu_update_dma(int64_t M, int64_t N, UpdSrc src,
                                                              float2* __restrict__ C, int64_t ldc, int nbm, int nbn) {
  __shared__ mg_dma::Stage s0, s1;
  constexpr int KG = kIB / 8, T = NSRC * (KG / 4);
  int tm, tn;
  mg_tile_order(blockIdx.x, nbm, nbn, tm, tn);
  const int64_t i0 = (int64_t)tm * BM, j0 = (int64_t)tn * BN;
  auto tile = [&](int t, mg_dma::Stage& st) {
    const int q = t / (KG / 4), g0 = (t % (KG / 4)) * 4;
    mg_dma::load_tile(st, src.A[q] + i0 * KG * U, src.B[q] + j0 * KG * U, KG, g0);
  };
  mg_dma::tile_update<T, LATE>(s0, s1, tile, C + i0 * ldc + j0, ldc, M - i0, N - j0);
}

// C is read late (by DMA, tile_update's LATE form) when its 16-byte units
// are aligned: standalone 12288^2 x 512, 1.59 against 1.70 ms; x 256, 0.99
// against 1.13 ms.
template <int NSRC>
int launch_dma(int64_t M, int64_t N, const UpdSrc& src, float2* C, int64_t ldc, hipStream_t s) {
  const int nbm = (int)((M + BM - 1) / BM), nbn = (int)((N + BN - 1) / BN);
  const bool late = (reinterpret_cast<uintptr_t>(C) & 15) == 0 && ldc % 2 == 0 && N % 2 == 0;
  if (late)
    hipLaunchKernelGGL((k_ldu_update_dma<NSRC, true>), dim3(nbm * nbn), dim3(mg_dma::NT), 0, s, M, N, src, C, ldc,
                       nbm, nbn);
  else
    hipLaunchKernelGGL((k_ldu_update_dma<NSRC, false>), dim3(nbm * nbn), dim3(mg_dma::NT), 0, s, M, N, src, C, ldc,
                       nbm, nbn);
  MG_HIP(hipGetLastError());
  return MG_OK;
}
}  // namespace dma

// C -= sum_s A_s B_s.  BF16x3: updates that give every CU a 128 x 128 tile
// run k_ldu_update_dma, smaller ones 64 x 128 tiles (two workgroups per CU;
// 12032^2 x 256 measured 1.07 ms against 1.21 ms with the register-staged
// 128 x 128 tile); BF16x6: 64 x 64 (its 64 x 128 tile's LDS would leave one
// workgroup per CU, and its 128 x 128 tile spills registers).
// whether a BF16x3 update of M x N takes the DMA-staged kernel
inline bool ldu_update_uses_dma(int64_t M, int64_t N) {
  return ((M + dma::BM - 1) / dma::BM) * ((N + dma::BN - 1) / dma::BN) >= (int64_t)mg::cu_count();
}

bool small_update_dma();
int launch_update_small(int64_t M, int64_t N, const UpdSrc& src, float2* C, int64_t ldc, hipStream_t s);

template <int NS>
int ldu_update(int64_t M, int64_t N, const UpdSrc& src, float2* C, int64_t ldc, hipStream_t s) {
  if (M <= 0 || N <= 0) return MG_OK;
  if (src.n < 1 || src.n > kMaxSrc) return MG_ERR_ARG;
  if constexpr (NS == 2) {
    static_assert(kIB % dma::BM == 0 && kIB % dma::BN == 0, "tiles within the padded 256-row blocks");
    if (ldu_update_uses_dma(M, N)) {
      switch (src.n) {
        case 1: return dma::launch_dma<1>(M, N, src, C, ldc, s);
        case 2: return dma::launch_dma<2>(M, N, src, C, ldc, s);
        case 3: return dma::launch_dma<3>(M, N, src, C, ldc, s);
        default: return dma::launch_dma<4>(M, N, src, C, ldc, s);
      }
    }
    if (small_update_dma()) return launch_update_small(M, N, src, C, ldc, s);
    return launch_update<NS, 64, 128, 2, 2>(M, N, src, C, ldc, s);
  } else {
    return launch_update<NS, 64, 64, 2, 2>(M, N, src, C, ldc, s);
  }
}

constexpr long long kSweepSpin = 1LL << 22;   // bound on one wait (~seconds)
constexpr int kSweepTickets = 8;              // ticket counters ahead of a sweep's block-row counters

// One block-triangular sweep of the LDU solve over the right-hand sides X
// (n x nrhs, row-major, ldx), in one launch:
//   forward (FWD):  x_i = b_i - sum_{j < i} Lh_ij x_j,   i = 0 .. nblk-1,
//   backward:       x_i = w_i - sum_{j > i} Uh_ij x_j,   i = nblk-1 .. 0,
// with b_i / w_i read from Bin (Bin may be X) and Lh_ij / Uh_ij from the
// factor's pre-split panels SP (SPL forward, SPUr backward).  One workgroup
// per BM x BN tile of X; it starts from -b_i and for each j in dependency
// order waits until every tile of block row j is final, then adds Lh_ij x_j
// (split-bf16 MFMA).  Finished tiles are published pre-split in XS (B-operand
// units, one column per right-hand side, np / 8 groups per column; rows past
// n as zeros) for the later block rows, and as complex float in X.
// ctr (zero on entry; next: the counters of the following sweep, cleared
// here) -- ctr[0] hands out tickets in start order and ctr[1 + i] counts the final
// tiles of block row i; tickets map to block rows in dependency order, so a
// workgroup only waits on workgroups that started before it and the sweep
// needs no co-residency.  XS is written with agent-scope stores and read with
// agent-scope loads (coherent across XCDs); each wave waits for its stores to
// complete before the barrier that precedes the counter increment.  A wait
// past the spin bound fills the tile with NaN and stops waiting, so every
// workgroup ends.
template <int NS, int BM, int BN, int WM, int WN, bool FWD>
__global__ __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(2))) void k_ldu_sweep(
    const u32x4* __restrict__ SP, int64_t np, const float2* Bin, int64_t ldbin, float2* X, int64_t ldx, u32x4* XS,
    int64_t n, int64_t nrhs, int nblk, int ct, int* ctr, int* next, int nnext) {
  typedef mg_ps::PTile<NS, BM, BN, WM, WN> Tile;
  constexpr int U = 2 * NS, KG = kIB / 8, KT = KG / 4;
  __shared__ typename Tile::S sm[2];
  __shared__ int s_ticket, s_ok;
  if (blockIdx.x == 0)   // the next sweep's whole counter set (that launch follows this one on the stream)
    for (int e = threadIdx.x; e < nnext; e += blockDim.x)
      __hip_atomic_store(next + e, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (threadIdx.x == 0) s_ticket = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int tpr = (kIB / BM) * ct;   // tiles per block row
  const int blk = s_ticket / tpr, w = s_ticket % tpr;
  const int i = FWD ? blk : nblk - 1 - blk;
  int* done = ctr + 1;
  const int64_t i0 = (int64_t)i * kIB + (int64_t)(w / ct) * BM, j0 = (int64_t)(w % ct) * BN;
  const int64_t M = n - i0, N = nrhs - j0;   // valid extents (M <= 0: a row tile past a short last block)
  const int64_t ldgx = np / 8;
  const SplitLayout lay{np, nblk, U};
  Tile tile;
  tile.init(Bin + i0 * ldbin + j0, ldbin, M, N, make_float2(-1.f, 0.f), M > 0);
  bool ok = true;
  const int steps = M > 0 ? (FWD ? i : nblk - 1 - i) : 0;
  // thread 0 polls block row j's counter (wait: until it is complete)
  auto poll_row = [&](int j, bool wait) {
    if (threadIdx.x == 0) {
      long long spins = 0;
      int st = 1;
      while (__hip_atomic_load(&done[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < tpr) {
        if (!wait) {
          st = 2;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kSweepSpin) {
          st = 0;
          break;
        }
      }
      s_ok = st;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return s_ok;   // 1: complete, 0: timed out, 2: not yet (poll only)
  };
  const u32x4* Aj = nullptr;
  const u32x4* Xj = nullptr;
  int64_t ldga = KG;
  int jcur = 0;
  auto set_step = [&](int st) {
    const int j = FWD ? st : nblk - 1 - st;
    jcur = j;
    if constexpr (FWD) {
      ldga = KG;
      Aj = SP + lay.off(j) + (i0 - (int64_t)(j + 1) * kIB) * KG * U;
    } else {
      ldga = lay.rows(i) / 8;
      Aj = SP + lay.off(i) + ((i0 - (int64_t)i * kIB) * ldga + (int64_t)(j - i - 1) * KG) * U;
    }
    Xj = XS + (j0 * ldgx + (int64_t)j * KG) * U;
  };
  // One loop over the k-tiles of all steps; the wait for block row j comes
  // before its first fetch.  When the next block row is not complete yet,
  // the current k-tile is multiplied before the wait.
  // Two LDS stages: k-tile i is multiplied from one while k-tile i + 1,
  // fetched during that product, is stashed into the other; one barrier per
  // k-tile.
  int st = 0, kt = 0, cur = 0;
  if (steps > 0) {
    set_step(0);
    ok = poll_row(jcur, true) == 1;
    if (ok) {
      tile.template pfetch<true>(Aj, ldga, Xj, ldgx, M, N, 0);
      tile.pstash(sm[0]);
    }
    __syncthreads();
  }
  while (ok && st < steps) {
    int nst = st, nkt = kt + 1;
    if (nkt == KT) {
      nkt = 0;
      ++nst;
    }
    const bool have_next = nst < steps;
    bool computed = false;
    if (have_next) {
      if (nkt == 0) {
        set_step(nst);
        int r = poll_row(jcur, false);
        if (r == 2) {   // not complete yet: multiply the current k-tile first
          tile.compute(sm[cur]);
          computed = true;
          r = poll_row(jcur, true);
        }
        ok = r == 1;
      }
      if (ok) tile.template pfetch<true>(Aj, ldga, Xj, ldgx, M, N, (int64_t)nkt * 4);
    }
    if (!computed) tile.compute(sm[cur]);
    if (ok && have_next) tile.pstash(sm[cur ^ 1]);
    __syncthreads();   // the next stage is complete, and this one's reads are done before it is refilled
    cur ^= 1;
    st = nst;
    kt = nkt;
  }
  if (!ok) {
    const float nan = __builtin_nanf("");
#pragma unroll
    for (int a = 0; a < Tile::TM; ++a)
#pragma unroll
      for (int b = 0; b < Tile::TN; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) tile.cr[a][b][r] = tile.ci[a][b][r] = nan;
  }
  tile.store_split_cols(XS + (j0 * ldgx + i0 / 8) * U, ldgx, M, N, make_float2(-1.f, 0.f));
  // every wave's stores complete before the barrier after which the row
  // counter is bumped (a workgroup fence emits no such wait on gfx950)
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_fetch_add(&done[i], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // the complex-float result is read only after the launch: stored after the signal
  if (M > 0)
    tile.store(X + i0 * ldx + j0, ldx, M, N, make_float2(-1.f, 0.f), make_float2(1.f, 0.f), false);
}

// The BF16x3 sweep with its operand k-tiles copied to LDS by DMA through
// three stages (k_ldu_sweep stages one k-tile ahead through registers, so
// every k-tile waited about one memory latency; the chain of block rows is
// made of such waits).  Tiles of RB rows (64 or 128) and 32 columns, waves
// of 32 x 16; the same counters and stores as k_ldu_sweep, tickets per XCD.
// x_j's units are read with agent-scope (sc1) DMA: other workgroups of the
// launch wrote them.  Every k-tile is fetched two ahead of its product; the
// counters of up to 8 block rows are read at once (a counter load waits for
// the wave's earlier loads, the DMA in flight included), when the first
// k-tile of a block row not yet known complete is to be fetched.
namespace sdma {
// RB rows per tile (64 or 128), 32 columns; 2 RB / 32 waves of 32 x 16
constexpr int BN = 32, UPR = 16;
template <int RB>
constexpr int nthreads() { return 4 * RB; }
// DMA instructions per wave and k-tile: A 16 rows (four rows each), B 512 / RB
// columns (four each)
template <int RB>
constexpr int dma_per_tile() { return 4 + 128 / RB; }
template <int RB>
struct Stage {
  u32x4 a[RB * UPR];
  u32x4 b[BN * UPR];
};
using mg_dma::gvp;
using mg_dma::kLdsWait;
using mg_dma::lvp;
using mg_dma::swz;
using mg_dma::vm_wait;

// One 16-byte-per-lane DMA (global_load_lds_dwordx4) of the wave's 64
// lanes into LDS at dst + 16 lane; SC1: agent-scope coherent read.  Written
// as inline assembly so that the compiler does not track it: it cannot tell
// which stage a DMA writes when the stage rotates in a run-time loop, and
// then waits for every DMA in flight before any LDS read, which undoes the
// prefetch.  The loop's own waits (vm_wait before the barrier) order the
// DMA and the reads; the instruction has no register results, and it reads
// its address registers when it issues.  One wait state separates the M0
// write from the instruction that reads it.
template <bool SC1>
__device__ inline void dma16(const u32x4* src, u32x4* dst) {
  const unsigned lds =
      __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)(__attribute__((address_space(3))) void*)dst);
  if constexpr (SC1)
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off sc1" ::"v"(src), "s"(lds)
                 : "memory", "m0");
  else
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(src), "s"(lds)
                 : "memory", "m0");
}

// acc += the k-tile in st; wave (wm, wn) holds rows 32 wm + [0, 32) and
// columns 16 wn + [0, 16) (the accumulator layout of PTile<2, RB, 32, RB / 32, 2>).
template <int RB>
__device__ inline void compute(const Stage<RB>& st, v4f (&cr)[2][1], v4f (&ci)[2][1]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1, li = lane & 15, lg = lane >> 4;
  const int sl = swz(li);
  bf16x8 ar[2][2], ai[2][2], br[2], bi[2], bn[2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int row = wm * 32 + a * 16 + li;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      ar[a][q] = mg_dma::unit(&st.a[row * UPR + ((lg * 4 + q) ^ sl)]);
      ai[a][q] = mg_dma::unit(&st.a[row * UPR + ((lg * 4 + 2 + q) ^ sl)]);
    }
  }
  const int col = wn * 16 + li;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    br[q] = mg_dma::unit(&st.b[col * UPR + ((lg * 4 + q) ^ sl)]);
    bi[q] = mg_dma::unit(&st.b[col * UPR + ((lg * 4 + 2 + q) ^ sl)]);
    bn[q] = mg_neg8(bi[q]);
  }
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q + p < 2; ++q) {
#pragma unroll
      for (int a = 0; a < 2; ++a) cr[a][0] = mg_mfma_bf16(ar[a][p], br[q], cr[a][0]);
#pragma unroll
      for (int a = 0; a < 2; ++a) ci[a][0] = mg_mfma_bf16(ar[a][p], bi[q], ci[a][0]);
#pragma unroll
      for (int a = 0; a < 2; ++a) cr[a][0] = mg_mfma_bf16(ai[a][p], bn[q], cr[a][0]);
#pragma unroll
      for (int a = 0; a < 2; ++a) ci[a][0] = mg_mfma_bf16(ai[a][p], br[q], ci[a][0]);
    }
}

template <int RB, bool FWD>
__global__ __launch_bounds__(nthreads<RB>()) void k_ldu_sweep_dma(const u32x4* __restrict__ SP, int64_t np, const float2* Bin,
                                                      int64_t ldbin, float2* X, int64_t ldx, u32x4* XS, int64_t n,
                                                      int64_t nrhs, int nblk, int ct, int* ctr, int* next,
                                                      int nnext) {
  constexpr int BM = RB, DPT = dma_per_tile<RB>();
  typedef Stage<RB> Stg;
  typedef mg_ps::PTile<2, BM, BN, BM / 32, 2> Tile;
  constexpr int U = 4, KG = kIB / 8, KT = KG / 4;
  // three stages, each named at every DMA and read (the loop is unrolled by three)
  __shared__ Stg sm0, sm1, sm2;
  __shared__ int s_ticket, s_cnt;
  if (blockIdx.x == 0)   // the next sweep's whole counter set (that launch follows this one on the stream)
    for (int e = threadIdx.x; e < nnext; e += blockDim.x)
      __hip_atomic_store(next + e, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // Tickets per XCD (workgroup b runs on XCD b mod 8): XCD x takes the row
  // tiles g = 4 blk + (row tile) with g mod 8 = x, block rows in sweep
  // order, each with its ct column tiles in a row, so that the column tiles
  // of a row tile share one L2 for the panel block they all read.  Every
  // class of workgroups hands out its tiles in dependency order and runs on
  // its own XCD, so a workgroup only waits on tiles that are done or running.
  const int xcd = blockIdx.x & 7;
  if (threadIdx.x == 0) s_ticket = __hip_atomic_fetch_add(ctr + xcd, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int tpr = (kIB / BM) * ct;   // tiles per block row
  const int g = xcd + 8 * (s_ticket / ct), ctile = s_ticket % ct;   // row tile (global), column tile
  if (g >= (kIB / BM) * nblk) return;                               // a workgroup past its XCD's tiles
  const int blk = g / (kIB / BM), rtile = g % (kIB / BM);
  const int i = FWD ? blk : nblk - 1 - blk;
  int* done = ctr + kSweepTickets;
  const int64_t i0 = (int64_t)i * kIB + (int64_t)rtile * BM, j0 = (int64_t)ctile * BN;
  const int64_t M = n - i0, N = nrhs - j0;   // valid extents (M <= 0: a row tile past a short last block)
  const int64_t ldgx = np / 8;
  const SplitLayout lay{np, nblk, U};
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  Tile tile;
  tile.init(Bin + i0 * ldbin + j0, ldbin, M, N, make_float2(-1.f, 0.f), M > 0);
  // the start values have landed before the loop: the compiler does not see
  // the loop's DMA, and would otherwise wait for all of it (vmcnt(0)) where
  // it first uses them
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(vm_wait(0));
  bool ok = true;
  const int steps = M > 0 ? (FWD ? i : nblk - 1 - i) : 0;
  const int Q = steps * KT;
  auto row_of_step = [&](int st) { return FWD ? st : nblk - 1 - st; };
  // Thread 0 reads the counters of the block rows of steps s .. s+7 (those
  // that exist) until the leading complete ones are at least one; returns
  // their count, -1 when the wait timed out.
  auto poll = [&](int s) {
    if (threadIdx.x == 0) {
      const int nb = min(8, steps - s);
      long long spins = 0;
      int c = 0;
      while (true) {
        int v[8];
#pragma unroll
        for (int b = 0; b < 8; ++b)
          v[b] = __hip_atomic_load(&done[row_of_step(min(s + b, steps - 1))], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        // every loaded value is used (a branch-free count): a load left in
        // flight would be waited for later by the compiler, which does not
        // see the DMA and so would wait for all of it
        c = 0;
        bool run = true;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          run = run && b < nb && v[b] >= tpr;
          c += run ? 1 : 0;
        }
        if (c > 0) break;
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kSweepSpin) {
          c = -1;
          break;
        }
      }
      s_cnt = c;
    }
    __builtin_amdgcn_s_waitcnt(kLdsWait);
    __builtin_amdgcn_s_barrier();
    const int c = s_cnt;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return c;
  };
  // k-tile q (step q / KT) into stage sg
  auto issue = [&](int q, Stg& sg) {
    const int st = q / KT, kt = q - st * KT, j = row_of_step(st);
    const u32x4* Aj;
    int64_t ldga;
    if constexpr (FWD) {
      ldga = KG;
      Aj = SP + lay.off(j) + (i0 - (int64_t)(j + 1) * kIB) * KG * U;
    } else {
      ldga = lay.rows(i) / 8;
      Aj = SP + lay.off(i) + ((i0 - (int64_t)i * kIB) * ldga + (int64_t)(j - i - 1) * KG) * U;
    }
    const u32x4* Xj = XS + (j0 * ldgx + (int64_t)j * KG) * U;
    const int64_t g0 = 4 * kt;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int r0 = wave * 16 + e * 4, r = r0 + (lane >> 4), u = (lane & 15) ^ swz(r);
      dma16<false>(Aj + ((int64_t)r * ldga + g0) * U + u, &sg.a[r0 * UPR]);
    }
#pragma unroll
    for (int e = 0; e < 128 / RB; ++e) {
      const int c0 = wave * (512 / RB) + e * 4, c = c0 + (lane >> 4), u = (lane & 15) ^ swz(c);
      const int64_t cc = c < N ? c : N - 1;
      dma16<true>(Xj + (cc * ldgx + g0) * U + u, &sg.b[c0 * UPR]);
    }
  };
  int known = 0;   // steps whose block rows are known complete
  // k-tile q into sg once its block row is known complete (reading counters,
  // and waiting, when it is not yet known)
  auto fetch = [&](int q, Stg& sg) {
    const int st = q / KT;
    if (st >= known) {
      const int c = poll(st);
      if (c <= 0) {
        ok = false;
        return;
      }
      known = st + c;
    }
    issue(q, sg);
  };
  // k-tile q from stage cur; k-tile q + 2 goes to n2, the stage of k-tile
  // q - 1 (free after the barrier).  Every k-tile is fetched two ahead of its
  // product and no other vector-memory load is in flight in the loop (the
  // counter loads are waited for where they are read), so "at most one
  // k-tile in flight" means k-tile q has landed.
  auto step = [&](int q, Stg& cur, Stg& n2) {
    asm volatile("" ::: "memory");   // the previous product's LDS reads stay above the barrier
    // k-tile q has landed when at most k-tile q + 1 is in flight
    if (q + 1 < Q) __builtin_amdgcn_s_waitcnt(vm_wait(DPT));
    else __builtin_amdgcn_s_waitcnt(vm_wait(0));
    // every wave's k-tile q has landed and its reads of stage n2 are done
    __builtin_amdgcn_s_waitcnt(kLdsWait);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");   // and this product's below it
    if (q + 2 < Q) fetch(q + 2, n2);
    compute<RB>(cur, tile.cr, tile.ci);
  };
  if (Q > 0) fetch(0, sm0);
  if (Q > 1 && ok) fetch(1, sm1);
  for (int q = 0; q < Q && ok; q += 3) {
    step(q, sm0, sm2);
    if (q + 1 < Q && ok) step(q + 1, sm1, sm0);
    if (q + 2 < Q && ok) step(q + 2, sm2, sm1);
  }
  __builtin_amdgcn_s_waitcnt(0);   // nothing in flight into LDS past this point
  __syncthreads();
  if (!ok) {
    const float nan = __builtin_nanf("");
#pragma unroll
    for (int a = 0; a < Tile::TM; ++a)
#pragma unroll
      for (int b = 0; b < Tile::TN; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) tile.cr[a][b][r] = tile.ci[a][b][r] = nan;
  }
  tile.store_split_cols(XS + (j0 * ldgx + i0 / 8) * U, ldgx, M, N, make_float2(-1.f, 0.f));
  // every wave's stores complete before the barrier after which the row
  // counter is bumped (a workgroup fence emits no such wait on gfx950)
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_fetch_add(&done[i], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (M > 0)
    tile.store(X + i0 * ldx + j0, ldx, M, N, make_float2(-1.f, 0.f), make_float2(1.f, 0.f), false);
}

// Trailing update C -= sum_s A_s B_s (as k_ldu_update, same operands) with
// the sweep's tiles and pipeline: RB x 32 tiles, operand k-tiles DMA'd two
// ahead through three stages; for the chain-side updates too small for a
// 128 x 128 tile per CU (the next diagonal block and the strips of the next
// super-step), which the register-staged tiles ran one memory latency per
// k-tile.  The operands hold whole tiles (padded panels), so the DMA reads
// need no bounds; C is loaded and stored within M x N.
template <int RB>
__global__ __launch_bounds__(nthreads<RB>()) void k_ldu_update_small(int64_t M, int64_t N, UpdSrc src,
                                                                    float2* __restrict__ C, int64_t ldc, int nbm,
                                                                    int nbn) {
  constexpr int BM = RB, DPT = dma_per_tile<RB>(), U = 4, KG = kIB / 8, KT = KG / 4;
  typedef Stage<RB> Stg;
  typedef mg_ps::PTile<2, BM, BN, BM / 32, 2> Tile;
  static_assert(kIB % BM == 0 && kIB % BN == 0, "tiles within the padded 256-row blocks");
  __shared__ Stg sm0, sm1, sm2;
  int tm, tn;
  mg_tile_order(blockIdx.x, nbm, nbn, tm, tn);
  const int64_t i0 = (int64_t)tm * BM, j0 = (int64_t)tn * BN;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  Tile tile;
  tile.init(C + i0 * ldc + j0, ldc, M - i0, N - j0, make_float2(-1.f, 0.f), true);
  // the start values have landed before the loop (the compiler does not see
  // the DMA; see dma16)
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(vm_wait(0));
  const int Q = src.n * KT;
  auto issue = [&](int q, Stg& sg) {
    const int sq = q / KT, kt = q - sq * KT;
    const u32x4* A = src.A[sq] + i0 * KG * U;
    const u32x4* B = src.B[sq] + j0 * KG * U;
    const int64_t g0 = 4 * kt;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int r0 = wave * 16 + e * 4, r = r0 + (lane >> 4), u = (lane & 15) ^ swz(r);
      dma16<false>(A + ((int64_t)r * KG + g0) * U + u, &sg.a[r0 * UPR]);
    }
#pragma unroll
    for (int e = 0; e < 128 / RB; ++e) {
      const int c0 = wave * (512 / RB) + e * 4, c = c0 + (lane >> 4), u = (lane & 15) ^ swz(c);
      dma16<false>(B + ((int64_t)c * KG + g0) * U + u, &sg.b[c0 * UPR]);
    }
  };
  // k-tile q from stage cur; k-tile q + 2 into n2 (the stage of k-tile q - 1)
  auto step = [&](int q, Stg& cur, Stg& n2) {
    asm volatile("" ::: "memory");
    if (q + 1 < Q) __builtin_amdgcn_s_waitcnt(vm_wait(DPT));
    else __builtin_amdgcn_s_waitcnt(vm_wait(0));
    __builtin_amdgcn_s_waitcnt(kLdsWait);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (q + 2 < Q) issue(q + 2, n2);
    compute<RB>(cur, tile.cr, tile.ci);
  };
  if (Q > 0) issue(0, sm0);
  if (Q > 1) issue(1, sm1);
  for (int q = 0; q < Q; q += 3) {
    step(q, sm0, sm2);
    if (q + 1 < Q) step(q + 1, sm1, sm0);
    if (q + 2 < Q) step(q + 2, sm2, sm1);
  }
  __builtin_amdgcn_s_waitcnt(0);
  tile.store(C + i0 * ldc + j0, ldc, M - i0, N - j0, make_float2(-1.f, 0.f), make_float2(1.f, 0.f), false);
}
}  // namespace sdma

// W = D^{-1} X block by block from the pre-split D inverses (DS: A-operand
// rows, 32 groups each) and the forward sweep's pre-split result (XS:
// B-operand columns, ldgx groups each); one workgroup per 64 x 64 tile of a
// block row, W complex float (n x nrhs, ldw).
template <int NS>
__global__ __launch_bounds__(256) void k_d_product(const u32x4* __restrict__ DS, const u32x4* __restrict__ XS,
                                                   int64_t ldgx, float2* __restrict__ W, int64_t ldw, int64_t n,
                                                   int64_t nrhs, int ct) {
  typedef mg_ps::PTile<NS, 64, 64, 2, 2> Tile;
  constexpr int U = 2 * NS, KG = kIB / 8;
  __shared__ typename Tile::S sm;
  const int tpb = (kIB / 64) * ct;
  const int blk = blockIdx.x / tpb, w = blockIdx.x - blk * tpb;
  const int64_t i0 = (int64_t)blk * kIB + (int64_t)(w / ct) * 64, j0 = (int64_t)(w % ct) * 64;
  const int64_t M = n - i0, N = nrhs - j0;
  if (M <= 0) return;
  Tile t;
  t.init(nullptr, 0, 0, 0, make_float2(0.f, 0.f), false);
  t.pmac(sm, DS + i0 * KG * U, KG, XS + (j0 * ldgx + (int64_t)blk * KG) * U, ldgx, M, N, KG);
  t.store(W + i0 * ldw + j0, ldw, M, N, make_float2(1.f, 0.f), make_float2(0.f, 0.f), false);
}

// Counter sets of the sweeps of one handle: two, used alternately; each
// sweep clears the other set for the next one (no zeroing launch).
struct SweepCtr {
  int* base;
  int64_t cap;   // ints per set
  int* parity;
  int* take() {
    int* c = base + cap * *parity;
    *parity ^= 1;
    return c;
  }
  int* other() const { return base + cap * *parity; }   // after take(): the next sweep's set
};

// Sweep launch: grid = block rows x tiles per block row.
template <int NS, int BM, int BN, int WM, int WN, bool FWD>
int launch_sweep(SweepCtr& sc, const u32x4* SP, int64_t np, const float2* Bin, int64_t ldbin, float2* X,
                 int64_t ldx, u32x4* XS, int64_t n, int64_t nrhs, hipStream_t s) {
  const int nblk = (int)((n + kIB - 1) / kIB);
  const int ct = (int)((nrhs + BN - 1) / BN);
  const int64_t grid = (int64_t)nblk * (kIB / BM) * ct;
  if (grid > 0x7fffffffLL) return MG_ERR_ARG;
  int* ctr = sc.take();
  hipLaunchKernelGGL((k_ldu_sweep<NS, BM, BN, WM, WN, FWD>), dim3((unsigned)grid), dim3(64 * WM * WN), 0, s, SP, np,
                     Bin, ldbin, X, ldx, XS, n, nrhs, nblk, ct, ctr, sc.other(), (int)sc.cap);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

// The LDU solve with pre-split panels: forward sweep, the block-diagonal
// product with the D inverses (split-bf16 MFMA from the pre-split inverses
// and the forward sweep's pre-split result), backward sweep.  On the
// 12288-row, 256-column solve with one LDS stage, 64 x 64 tiles measured
// 1.15 ms per sweep against 1.4 ms for 64 x 128 and 1.2-1.5 ms for 32 x 64,
// 64 x 32 and 32 x 32 (the sweep's time is its chain of 48 dependent block
// rows, not its MFMA work); with two LDS stages 64 x 64 took 1.08-1.10 ms (one
// workgroup per CU) and 64 x 32 1.05-1.07 ms (63 KB: two per CU; the row's
// last product 8.7 us against 10.7), 0.7 ms less per refined solve.
// The chain-side BF16x3 updates (C too small for a 128 x 128 tile per CU):
// 64 x 32 tiles staged by DMA (sdma::k_ldu_update_small); MG_SMALL_UPDATE=0
// selects the register-staged 64 x 128 tiles (A/B timing).
bool small_update_dma() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("MG_SMALL_UPDATE");
    v = (e && e[0] == '0') ? 0 : 1;
  }
  return v == 1;
}

int launch_update_small(int64_t M, int64_t N, const UpdSrc& src, float2* C, int64_t ldc, hipStream_t s) {
  constexpr int RB = 64;
  const int nbm = (int)((M + RB - 1) / RB), nbn = (int)((N + sdma::BN - 1) / sdma::BN);
  hipLaunchKernelGGL((sdma::k_ldu_update_small<RB>), dim3(nbm * nbn), dim3(sdma::nthreads<RB>()), 0, s, M, N, src,
                     C, ldc, nbm, nbn);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

// DMA-staged BF16x3 sweep launch: RB-row tiles
template <int RB, bool FWD>
int launch_sweep_dma(SweepCtr& sc, const u32x4* SP, int64_t np, const float2* Bin, int64_t ldbin, float2* X,
                     int64_t ldx, u32x4* XS, int64_t n, int64_t nrhs, hipStream_t s) {
  const int nblk = (int)((n + kIB - 1) / kIB);
  const int ct = (int)((nrhs + sdma::BN - 1) / sdma::BN);
  // eight classes of workgroups (one per XCD), each with the column tiles of
  // every eighth row tile
  const int64_t row_tiles = (int64_t)nblk * (kIB / RB);
  const int64_t grid = 8 * ((row_tiles + 7) / 8) * ct;
  if (grid > 0x7fffffffLL) return MG_ERR_ARG;
  int* ctr = sc.take();
  hipLaunchKernelGGL((sdma::k_ldu_sweep_dma<RB, FWD>), dim3((unsigned)grid), dim3(sdma::nthreads<RB>()), 0, s, SP,
                     np, Bin, ldbin, X, ldx, XS, n, nrhs, nblk, ct, ctr, sc.other(), (int)sc.cap);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

// The BF16x3 sweeps' tiles: 128 rows (one workgroup of eight waves per CU;
// the published rows x_j are read by half as many row tiles as with 64, two
// workgroups per CU): standalone 1.52 against 1.61 ms per preconditioner
// application.  MG_SWEEP_DMA: 64 or 128 rows, 0 the register-staged sweeps
// (A/B timing).
inline int sweep_stages() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("MG_SWEEP_DMA");
    v = e ? atoi(e) : 128;
    if (v != 0 && v != 64 && v != 128) v = 128;
  }
  return v;
}

template <int NS>
int ldu_sweeps(SweepCtr& ctr, const u32x4* SPL, const u32x4* SPUr, const u32x4* DS, int64_t np, float2* B,
               int64_t ldb, int64_t nrhs, float2* W, u32x4* XS, int64_t n, hipStream_t s) {
  constexpr int BN = 32;
  const int64_t nblk = (n + kIB - 1) / kIB;
  const int ct = (int)((nrhs + 63) / 64);
  const int dma = NS == 2 ? sweep_stages() : 0;
  if (dma == 64) MG_TRY((launch_sweep_dma<64, true>(ctr, SPL, np, B, ldb, B, ldb, XS, n, nrhs, s)));
  else if (dma == 128) MG_TRY((launch_sweep_dma<128, true>(ctr, SPL, np, B, ldb, B, ldb, XS, n, nrhs, s)));
  else MG_TRY((launch_sweep<NS, 64, BN, 2, 2, true>(ctr, SPL, np, B, ldb, B, ldb, XS, n, nrhs, s)));
  hipLaunchKernelGGL(k_d_product<NS>, dim3((unsigned)(nblk * (kIB / 64) * ct)), dim3(256), 0, s, DS, XS, np / 8, W,
                     nrhs, n, nrhs, ct);
  MG_HIP(hipGetLastError());
  if (dma == 64) return launch_sweep_dma<64, false>(ctr, SPUr, np, W, nrhs, B, ldb, XS, n, nrhs, s);
  if (dma == 128) return launch_sweep_dma<128, false>(ctr, SPUr, np, W, nrhs, B, ldb, XS, n, nrhs, s);
  return launch_sweep<NS, 64, BN, 2, 2, false>(ctr, SPUr, np, W, nrhs, B, ldb, XS, n, nrhs, s);
}

// R[i][j] = 2 delta_ij for i, j < h (row stride kIB)
__global__ void k_two_eye(float2* R, int h) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= h * kIB) return;
  const int i = e / kIB, j = e % kIB;
  R[e] = make_float2(i == j ? 2.f : 0.f, 0.f);
}

}  // namespace

namespace mg {

int64_t ldu_dinv_elems(int64_t n) { return ((n + kIB - 1) / kIB) * (int64_t)kIB * kIB; }

namespace {
int64_t padded(int64_t n) { return ((n + kIB - 1) / kIB) * kIB; }
// complex-float elements (8 bytes) of the pre-split panels of a split precision
int64_t split_elems(int ns, int64_t n) {
  if (ns == 0) return 0;
  const SplitLayout lay{padded(n), (n + kIB - 1) / kIB, 2 * ns};
  // SPL and SPUr, then DS (the D inverses' rows, 32 groups each), two float2 per 16-byte unit
  return (2 * lay.total() + padded(n) * (kIB / 8) * 2 * ns) * 2;
}
// the pre-split copies of A12, one per step for four steps in flight (widest split, 6 units)
int64_t a12_split_elems(int64_t n) { return 4 * padded(n) * (kIB / 8) * 6 * 2; }
}  // namespace

int64_t ldu_factor_elems(int ns, int64_t n) { return ldu_dinv_elems(n) + split_elems(ns, n); }
int64_t ldu_work_elems(int64_t n) { return 4 * (int64_t)kIB * n + 2 * (int64_t)kIB * kIB + a12_split_elems(n); }
int64_t ldu_solve_work_elems(int ns, int64_t n, int64_t nrhs) {
  const int64_t w = n * nrhs;
  if (ns == 0) return w;
  // W, then (16-byte aligned) the pre-split right-hand sides: nrhs columns of np / 8 groups
  return ((w + 1) & ~(int64_t)1) + nrhs * (padded(n) / 8) * (2 * ns) * 2;
}

int getrf_ldu(Workspace& ws, void* A, int64_t n, int64_t lda, void* dinv, void* work, int* info, hipStream_t s,
              GemmFn gemm, GemmFn update, bool newton, int split_ns) {
  if (n < 0 || lda < n) return MG_ERR_ARG;
  MG_TRY(workspace_init(ws));
  hipLaunchKernelGGL(k_zero_info, dim3(1), dim3(1), 0, s, info);
  if (n == 0) return MG_OK;
  static_assert(kIB == kInvBlock, "diagonal block order");
  float2* Ap = reinterpret_cast<float2*>(A);
  float2* D = reinterpret_cast<float2*>(dinv);
  float2* W = reinterpret_cast<float2*>(work);
  const GemmFn g = gemm ? gemm : &cgemm<float>;
  const GemmFn gt = update ? update : (gemm ? gemm : &cgemm_trail<float>);
  const int64_t nb = kIB, nblk = (n + nb - 1) / nb;
  // scratch copies of A12 (h x r) and A21 (r x h) per step, double-buffered by step parity
  auto s12 = [&](int64_t k) { return W + (k & 1) * 2 * nb * n; };
  auto s21 = [&](int64_t k) { return W + (k & 1) * 2 * nb * n + nb * n; };
  // step k's inverse and panels, on stream st
  // With FP32-level products the BF16x3 block inverse would limit the
  // factor's accuracy (newton): one Newton-Schulz step X <- X (2I - S X) in
  // FP32 squares its error.  X0 and R live after the panel copies in the work
  // buffer (panels run one at a time).
  float2* X0 = W + 4 * nb * n;
  float2* R = X0 + nb * nb;
  // split precisions: the pre-split panels (factor data after the D
  // inverses) and the pre-split A12 of the current step (work, after R)
  const int U = 2 * split_ns;
  const SplitLayout lay{padded(n), nblk, U};
  u32x4* SPL = split_ns ? reinterpret_cast<u32x4*>(D + ldu_dinv_elems(n)) : nullptr;
  u32x4* SPUr = split_ns ? SPL + lay.total() : nullptr;
  u32x4* A12s = reinterpret_cast<u32x4*>(R + nb * nb);
  auto spa12 = [&](int64_t k) { return A12s + (k & 3) * lay.np * (kIB / 8) * U; };
  // step k's panels of a split precision, straight from A and D_k: A12
  // pre-split as the trailing update's B operand, Uh = D_k A12 and
  // Lh = A21 D_k as GEMMs whose epilogues write the pre-split panels (A is
  // not written, so no copies are needed).  A12's split and then Uh, which
  // the solves alone read, are formed on the fourth stream beside Lh (A12
  // and D_k are final here and never written again); the chain waits for
  // the split only (evH), the factorisation's end joins that stream.
  auto split_panels = [&](int64_t k, hipStream_t st) -> int {
    const int64_t r0 = k * nb, r1 = r0 + nb, r = n - r1, rk = lay.rows(k);
    const float2* Dk = D + k * nb * nb;
    const float2* A12 = Ap + r0 * lda + r1;
    const float2* A21 = Ap + r1 * lda + r0;
    hipStream_t su = ws.s4 ? ws.s4 : st;
    if (su != st) {
      MG_HIP(hipEventRecord(ws.evF, st));
      MG_HIP(hipStreamWaitEvent(su, ws.evF, 0));
    }
    if (split_ns == 2) MG_TRY(pack_cols<2>(A12, lda, nb, r, rk, kIB / 8, spa12(k), su));
    else MG_TRY(pack_cols<3>(A12, lda, nb, r, rk, kIB / 8, spa12(k), su));
    if (su != st) MG_HIP(hipEventRecord(ws.evH, su));
    if (split_ns == 2) {
      MG_TRY(panel_split<2>(nb, r, nb, Dk, nb, A12, lda, nb, rk, SPUr + lay.off(k), rk / 8, su));
      MG_TRY(panel_split<2>(r, nb, nb, A21, lda, Dk, nb, rk, nb, SPL + lay.off(k), kIB / 8, st));
    } else {
      MG_TRY(panel_split<3>(nb, r, nb, Dk, nb, A12, lda, nb, rk, SPUr + lay.off(k), rk / 8, su));
      MG_TRY(panel_split<3>(r, nb, nb, A21, lda, Dk, nb, rk, nb, SPL + lay.off(k), kIB / 8, st));
    }
    if (su != st) MG_HIP(hipStreamWaitEvent(st, ws.evH, 0));
    return MG_OK;
  };
  auto inverse = [&](int64_t k, hipStream_t st) -> int {
    const int64_t r0 = k * nb, h = (n - r0 < nb) ? n - r0 : nb;
    float2* Dk = D + k * nb * nb;
    const float2* S = Ap + r0 * lda + r0;
    // the two control sets alternate: this launch's was cleared by the
    // previous one, which this one clears in turn (launches are ordered)
    int* ctl = ws.inv_ctl + 64 * ws.inv_parity;
    int* other = ws.inv_ctl + 64 * (1 - ws.inv_parity);
    hipLaunchKernelGGL(k_block_inverse, dim3(kSlabs), dim3(kMwThreads), 0, st, S, lda, (int)h, newton ? X0 : Dk,
                       info, r0, ws.inv_exch, reinterpret_cast<unsigned*>(ctl), ctl + 32, other);
    MG_HIP(hipGetLastError());
    ws.inv_parity ^= 1;
    if (newton) {
      hipLaunchKernelGGL(k_two_eye, dim3((unsigned)((h * nb + 255) / 256)), dim3(256), 0, st, R, (int)h);
      MG_HIP(hipGetLastError());
      MG_TRY(cgemm<float>(h, h, h, cd{-1, 0}, S, lda, X0, nb, cd{1, 0}, R, nb, st));    // R = 2I - S X0
      MG_TRY(cgemm<float>(h, h, h, cd{1, 0}, X0, nb, R, nb, cd{0, 0}, Dk, nb, st));     // D_k = X0 R
    }
    return MG_OK;
  };
  // step k's panels from its inverse (and, split precisions, their pre-split copies)
  auto panel_rest = [&](int64_t k, hipStream_t st) -> int {
    const int64_t r0 = k * nb, h = (n - r0 < nb) ? n - r0 : nb, r1 = r0 + h, r = n - r1;
    if (r == 0) return MG_OK;
    if (split_ns) return split_panels(k, st);
    const float2* Dk = D + k * nb * nb;
    float2 *S12 = s12(k), *S21 = s21(k);
    MG_HIP(hipMemcpy2DAsync(S12, r * sizeof(float2), Ap + r0 * lda + r1, lda * sizeof(float2), r * sizeof(float2),
                            h, hipMemcpyDeviceToDevice, st));
    MG_HIP(hipMemcpy2DAsync(S21, h * sizeof(float2), Ap + r1 * lda + r0, lda * sizeof(float2), h * sizeof(float2),
                            r, hipMemcpyDeviceToDevice, st));
    MG_TRY(g(h, r, h, cd{1, 0}, Dk, nb, S12, r, cd{0, 0}, Ap + r0 * lda + r1, lda, st));   // Uh = S^-1 A12
    MG_TRY(g(r, h, h, cd{1, 0}, S21, h, Dk, nb, cd{0, 0}, Ap + r1 * lda + r0, lda, st));   // Lh = A21 S^-1
    return MG_OK;
  };
  // A22[rows i0.., cols j0..j0+nc] -= A21 S^-1 A12 for step k (A21 from the scratch copy, Uh in place)
  auto trail = [&](int64_t k, int64_t i0, int64_t mr, int64_t j0, int64_t nc, hipStream_t st) -> int {
    if (mr <= 0 || nc <= 0) return MG_OK;
    const int64_t r0 = k * nb, h = (n - r0 < nb) ? n - r0 : nb, r1 = r0 + h;
    return timed_update(ws, gt, mr, nc, h, cd{-1, 0}, s21(k) + (i0 - r1) * h, h, Ap + r0 * lda + j0, lda, cd{1, 0},
                        Ap + i0 * lda + j0, lda, st);
  };
  // split precisions: A22[rows i0.., cols j0..] -= sum over steps k .. k+cnt-1 of Lh A12, pre-split
  auto trail_ps = [&](int64_t k, int cnt, int64_t i0, int64_t mr, int64_t j0, int64_t nc, hipStream_t st) -> int {
    if (mr <= 0 || nc <= 0) return MG_OK;
    constexpr int64_t KG = kIB / 8;
    UpdSrc src;
    src.n = cnt;
    for (int q = 0; q < kMaxSrc; ++q) {
      const int64_t kq = k + (q < cnt ? q : 0), r1 = (kq + 1) * nb;
      src.A[q] = SPL + lay.off(kq) + (i0 - r1) * KG * U;
      src.B[q] = spa12(kq) + (j0 - r1) * KG * U;
    }
    const double flops = 8.0 * (double)mr * (double)nc * (double)(cnt * nb);
    return timed_launch(
        ws, flops, st,
        [&] {
          return split_ns == 2 ? ldu_update<2>(mr, nc, src, Ap + i0 * lda + j0, lda, st)
                               : ldu_update<3>(mr, nc, src, Ap + i0 * lda + j0, lda, st);
        },
        (split_ns == 2 && ldu_update_uses_dma(mr, nc)) ? 1 : 0);
  };
  // split precisions: the D inverses pre-split as A operands (DS, after
  // SPUr) for the solves' block-diagonal products, once all are final on s
  // (full blocks in one launch; a short last block with its h x h inverse
  // only, zeros around it: the Newton step writes no padding)
  auto pack_d = [&]() -> int {
    if (!split_ns) return MG_OK;
    if (ws.s4) {   // the Uh panels
      MG_HIP(hipEventRecord(ws.evG, ws.s4));
      MG_HIP(hipStreamWaitEvent(s, ws.evG, 0));
    }
    u32x4* DS = SPUr + lay.total();
    const int64_t full = n / nb, h = n - full * nb;
    auto pack = [&](const float2* src, int64_t rows, int64_t K, int64_t rows_pad, u32x4* out) {
      const dim3 grid((unsigned)((rows_pad * (kIB / 8) + 255) / 256));
      if (split_ns == 2)
        hipLaunchKernelGGL(k_pack_rows<2>, grid, dim3(256), 0, s, src, nb, rows, K, rows_pad, kIB / 8, out);
      else
        hipLaunchKernelGGL(k_pack_rows<3>, grid, dim3(256), 0, s, src, nb, rows, K, rows_pad, kIB / 8, out);
    };
    if (full > 0) pack(D, full * nb, nb, full * nb, DS);
    if (h > 0) pack(D + full * nb * nb, h, h, nb, DS + full * nb * (kIB / 8) * U);
    MG_HIP(hipGetLastError());
    return MG_OK;
  };
  if (!split_ns || !ws.s2 || !ws.s3) {
    MG_TRY(inverse(0, s));
    MG_TRY(panel_rest(0, s));
    // Look-ahead of depth one (with a second stream): block row and column
    // k+1 are updated first, step k+1's inverse and panels run on the second
    // stream while the rest of step k's trailing update runs here.
    for (int64_t k = 0; k < nblk; ++k) {
      const int64_t r0 = k * nb, h = (n - r0 < nb) ? n - r0 : nb, r1 = r0 + h, r = n - r1;
      if (r == 0) break;
      const int64_t h2 = (r < nb) ? r : nb;   // next block's order
      auto upd = [&](int64_t i0, int64_t mr, int64_t j0, int64_t nc, hipStream_t st) {
        return split_ns ? trail_ps(k, 1, i0, mr, j0, nc, st) : trail(k, i0, mr, j0, nc, st);
      };
      MG_TRY(upd(r1, h2, r1, r, s));
      MG_TRY(upd(r1 + h2, r - h2, r1, h2, s));
      if (ws.s2) {
        MG_HIP(hipEventRecord(ws.evA, s));
        MG_HIP(hipStreamWaitEvent(ws.s2, ws.evA, 0));
        MG_TRY(inverse(k + 1, ws.s2));
        MG_TRY(panel_rest(k + 1, ws.s2));
        MG_HIP(hipEventRecord(ws.evB, ws.s2));
        MG_TRY(upd(r1 + h2, r - h2, r1 + h2, r - h2, s));
        MG_HIP(hipStreamWaitEvent(s, ws.evB, 0));
      } else {
        MG_TRY(upd(r1 + h2, r - h2, r1 + h2, r - h2, s));
        MG_TRY(inverse(k + 1, s));
        MG_TRY(panel_rest(k + 1, s));
      }
    }
    return pack_d();
  }
  // Split precisions, in super-steps of two blocks (k, k+1).  The second
  // stream carries the critical path: the inverse and panels of block k, the
  // update of the next diagonal block S_{k+1} by step k, the inverse and
  // panels of block k+1; the rest of block row and column k+1 is updated on
  // the third stream during that inversion.  The caller's stream updates the
  // next super-step's block rows and columns (k+2, k+3) by both steps at once
  // (k = 512 products) while the second stream updates S_{k+2} alone and
  // inverts it; the next super-step's panels follow, and the rest of the
  // trailing matrix is updated beside them, again by both steps at once: one
  // read and write of the trailing matrix per two steps.  Concurrent launches
  // touch disjoint blocks.  evA: a super-step's panels done; evB: the next
  // super-step's block rows and columns done; evC / evD: step k's panels /
  // block row and column k+1 done (third stream); evE: the next diagonal
  // block S_{k+4} updated by the super-step (third stream).
  auto chain_tail = [&](int64_t k, hipStream_t st) -> int {   // after inverse(k)
    MG_TRY(panel_rest(k, st));
    if (k + 1 >= nblk) return MG_OK;
    const int64_t r1 = (k + 1) * nb, h2 = (n - r1 < nb) ? n - r1 : nb, r2 = r1 + h2;
    MG_TRY(trail_ps(k, 1, r1, h2, r1, h2, st));           // S_{k+1}
    MG_HIP(hipEventRecord(ws.evC, st));
    MG_HIP(hipStreamWaitEvent(ws.s3, ws.evC, 0));
    MG_TRY(trail_ps(k, 1, r1, h2, r2, n - r2, ws.s3));   // block row k+1 right of S_{k+1}
    MG_TRY(trail_ps(k, 1, r2, n - r2, r1, h2, ws.s3));   // block column k+1 below it
    MG_HIP(hipEventRecord(ws.evD, ws.s3));
    MG_TRY(inverse(k + 1, st));
    MG_HIP(hipStreamWaitEvent(st, ws.evD, 0));
    return panel_rest(k + 1, st);
  };
  MG_TRY(inverse(0, s));
  MG_TRY(chain_tail(0, s));
  MG_HIP(hipEventRecord(ws.evA, s));
  for (int64_t k = 0; k + 2 < nblk; k += 2) {
    const int64_t b2 = (k + 2) * nb, w = (n - b2 < 2 * nb) ? n - b2 : 2 * nb, b4 = b2 + w;
    const int64_t h3 = (n - b2 < nb) ? n - b2 : nb, b3 = b2 + h3;   // block k+2
    const int64_t h5 = (n - b4 < nb) ? n - b4 : nb, b5 = b4 + h5;   // block k+4
    // S_{k+2}: steps k and k+1 (steps k-2 and k-1, whose rest left it out,
    // were applied on the third stream during the previous rest)
    MG_HIP(hipStreamWaitEvent(ws.s2, ws.evA, 0));
    if (k >= 2) MG_HIP(hipStreamWaitEvent(ws.s2, ws.evE, 0));
    MG_TRY(trail_ps(k, 2, b2, h3, b2, h3, ws.s2));
    MG_TRY(inverse(k + 2, ws.s2));
    MG_TRY(trail_ps(k, 2, b2, h3, b3, n - b3, s));       // block row k+2 right of S_{k+2}
    MG_TRY(trail_ps(k, 2, b3, b4 - b3, b2, n - b2, s));   // block row k+3
    MG_TRY(trail_ps(k, 2, b4, n - b4, b2, w, s));         // block columns k+2, k+3 below them
    MG_HIP(hipEventRecord(ws.evB, s));
    MG_HIP(hipStreamWaitEvent(ws.s2, ws.evB, 0));
    MG_TRY(chain_tail(k + 2, ws.s2));
    MG_HIP(hipEventRecord(ws.evA, ws.s2));
    // the rest, without S_{k+4}, whose update by these two steps runs beside
    // it on the third stream (evE) so that the next super-step's chain only
    // adds its own two steps there
    MG_TRY(trail_ps(k, 2, b4, h5, b5, n - b5, s));
    MG_TRY(trail_ps(k, 2, b5, n - b5, b4, n - b4, s));
    MG_HIP(hipStreamWaitEvent(ws.s3, ws.evB, 0));
    MG_TRY(trail_ps(k, 2, b4, h5, b4, h5, ws.s3));
    MG_HIP(hipEventRecord(ws.evE, ws.s3));
    MG_HIP(hipStreamWaitEvent(s, ws.evA, 0));
    MG_HIP(hipStreamWaitEvent(s, ws.evE, 0));
  }
  return pack_d();
}

int getrs_ldu(Workspace* ws, const void* LDU, int64_t n, int64_t lda, const void* dinv, void* B, int64_t ldb,
              int64_t nrhs, void* work, hipStream_t s, GemmFn gemm, int sweep_ns) {
  if (n < 0 || nrhs < 0 || lda < n || ldb < nrhs) return MG_ERR_ARG;
  if (n == 0 || nrhs == 0) return MG_OK;
  const GemmFn g = gemm ? gemm : &cgemm<float>;
  const float2* Ap = reinterpret_cast<const float2*>(LDU);
  const float2* D = reinterpret_cast<const float2*>(dinv);
  float2* Bp = reinterpret_cast<float2*>(B);
  float2* W = reinterpret_cast<float2*>(work);
  const int64_t nb = kIB, nblk = (n + nb - 1) / nb, full = n / nb;
  if (sweep_ns == 2 || sweep_ns == 3) {
    if (!ws) return MG_ERR_ARG;
    if (ws->sweep_cap < nblk + kSweepTickets) {
      // grows rarely (a larger system than any before on this handle)
      MG_HIP(hipStreamSynchronize(s));
      if (ws->sweep) MG_HIP(hipFree(ws->sweep));
      ws->sweep = nullptr;
      ws->sweep_cap = 0;
      const int64_t cap = ((nblk + kSweepTickets + 255) / 256) * 256;
      MG_HIP(hipMalloc(&ws->sweep, sizeof(int) * 2 * cap));
      hipLaunchKernelGGL(k_zero_ints, dim3((unsigned)((2 * cap + 255) / 256)), dim3(256), 0, s, ws->sweep,
                         (int)(2 * cap));
      MG_HIP(hipGetLastError());
      ws->sweep_cap = cap;
      ws->sweep_parity = 0;
    }
    SweepCtr ctr{ws->sweep, ws->sweep_cap, &ws->sweep_parity};
    const SplitLayout lay{padded(n), nblk, 2 * sweep_ns};
    const u32x4* SPL = reinterpret_cast<const u32x4*>(D + ldu_dinv_elems(n));
    const u32x4* SPUr = SPL + lay.total();
    const u32x4* DS = SPUr + lay.total();
    u32x4* XS = reinterpret_cast<u32x4*>(W + ((n * nrhs + 1) & ~(int64_t)1));
    if (sweep_ns == 2) return ldu_sweeps<2>(ctr, SPL, SPUr, DS, lay.np, Bp, ldb, nrhs, W, XS, n, s);
    return ldu_sweeps<3>(ctr, SPL, SPUr, DS, lay.np, Bp, ldb, nrhs, W, XS, n, s);
  }
  // forward: B[r1:] -= Lh[r1:, block k] B[block k]
  for (int64_t k = 0; k + 1 < nblk; ++k) {
    const int64_t r0 = k * nb, r1 = r0 + nb;
    MG_TRY(g(n - r1, nrhs, nb, cd{-1, 0}, Ap + r1 * lda + r0, lda, Bp + r0 * ldb, ldb, cd{1, 0}, Bp + r1 * ldb, ldb,
             s));
  }
  // W = D^{-1} B, block by block (FP32 products)
  if (full > 0)
    MG_TRY(cgemm_batched<float>(nb, nrhs, nb, cd{1, 0}, D, nb, nb * nb, Bp, ldb, nb * ldb, cd{0, 0}, W, nrhs,
                                nb * nrhs, full, s));
  if (full < nblk) {
    const int64_t r0 = full * nb, h = n - r0;
    MG_TRY(cgemm<float>(h, nrhs, h, cd{1, 0}, D + full * nb * nb, nb, Bp + r0 * ldb, ldb, cd{0, 0}, W + r0 * nrhs,
                        nrhs, s));
  }
  // backward: W[:r0] -= Uh[:r0, block k] W[block k]
  for (int64_t k = nblk - 1; k > 0; --k) {
    const int64_t r0 = k * nb, h = (n - r0 < nb) ? n - r0 : nb;
    MG_TRY(g(r0, nrhs, h, cd{-1, 0}, Ap + r0, lda, W + r0 * nrhs, nrhs, cd{1, 0}, W, nrhs, s));
  }
  MG_HIP(hipMemcpy2DAsync(Bp, ldb * sizeof(float2), W, nrhs * sizeof(float2), nrhs * sizeof(float2), n,
                          hipMemcpyDeviceToDevice, s));
  return MG_OK;
}

}  // namespace mg
