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
#include "lu.h"
#include "mg_internal.h"
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
  float2 pf[kPW][kPW + 1];            // P^{-1}
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
//   wave 0 of every workgroup inverts P in registers (FP32, lane = column,
//   rows through cross-lane shuffles);
//   R = P^{-1} X_p (the slab's pivot rows; R = P^{-1} on slab p) becomes the
//   slab's pivot rows, and every other row gets X_i <- X_i - C_i R (slab p:
//   X_i <- -C_i P^{-1}) as BF16x3 MFMA products.
// One barrier per panel; zero pivots are reported in info (first, 1-based).
__global__ __launch_bounds__(kMwThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_block_inverse(const float2* __restrict__ A, int64_t lda, int h,
                                                              float2* __restrict__ X, int* info, int64_t row0,
                                                              float2* exch, unsigned* bar, int* err) {
  __shared__ InvSmem sm;
  const int q = blockIdx.x;
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
    // (1) P^{-1}: lane (j = lane % 32, half hh) holds P[16 hh + t][j], t = 0..15
    if (wave == 0) {
      const int j = lane & 31, hh = lane >> 5;
      float2 v[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) v[t] = mg_sg::coherent_load(&C[(kPW * p + 16 * hh + t) * kPW + j]);
#pragma unroll
      for (int st = 0; st < kPW; ++st) {
        const int hs = st >> 4, ts = st & 15;
        const float px = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[ts].x), hs * 32 + st));
        const float py = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[ts].y), hs * 32 + st));
        if (q == 0 && lane == 0 && px == 0.f && py == 0.f && kPW * p + st < h)
          atomicCAS(info, 0, (int)(row0 + kPW * p + st + 1));
        const float rd = __builtin_amdgcn_rcpf(px * px + py * py);
        const float2 inv = make_float2(px * rd, -py * rd);
        const float2 rowv = make_float2(__shfl(v[ts].x, hs * 32 + j), __shfl(v[ts].y, hs * 32 + j));
        const float2 u = (j == st) ? inv : cmulf(rowv, inv);   // row st of the result
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          const float2 cv = make_float2(__shfl(v[t].x, hh * 32 + st), __shfl(v[t].y, hh * 32 + st));
          const float2 tq = cmulf(cv, (j == st) ? inv : u);
          const int i = 16 * hh + t;
          v[t] = (i == st) ? u : (j == st) ? make_float2(-tq.x, -tq.y) : make_float2(v[t].x - tq.x, v[t].y - tq.y);
        }
      }
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const int i = 16 * hh + t;
        sm.pf[i][j] = v[t];
        split_planes(v[t], &sm.pa[0][i][j], &sm.pa[1][i][j], &sm.pa[2][i][j], &sm.pa[3][i][j]);
      }
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

constexpr long long kSweepSpin = 1LL << 22;   // bound on one wait (~seconds)

// One block-triangular sweep of the LDU solve over the right-hand sides X
// (n x nrhs, row-major, ldx), in one launch:
//   forward (FWD):  x_i = b_i - sum_{j < i} Lh_ij x_j,   i = 0 .. nblk-1,
//   backward:       x_i = w_i - sum_{j > i} Uh_ij x_j,   i = nblk-1 .. 0,
// Lh_ij / Uh_ij the 256 x 256 blocks of LDU (row-major, lda) and b_i / w_i
// read from Bin (Bin may be X).  One workgroup per BM x BN tile of X; it
// starts from -b_i, and for each j in dependency order waits until every
// tile of block row j is final, then adds Lh_ij x_j (split-bf16 MFMA).  The
// workgroup holds KS thread groups of WM x WN waves, each with its own
// accumulators and LDS planes; group g takes k-tiles g, g + KS, ... of every
// block, which divides the chain of dependent k-tiles per block by KS, and
// the groups' sums are added through LDS at the end.
// ctr[0] hands out tickets in start order and ctr[1 + i] counts the final
// tiles of block row i; tickets map to block rows in dependency order, so a
// workgroup only waits on workgroups that started before it and the sweep
// needs no co-residency.  Final tiles are written with agent-scope stores and
// read with agent-scope loads (coherent across XCDs); each wave waits for its
// stores to complete before the barrier that precedes the counter increment.
// A wait that exceeds the spin bound fills its tile with NaN and stops
// waiting, so every workgroup ends.
template <int NS, int BM, int BN, int WM, int WN, int KS, bool FWD>
__global__ __launch_bounds__(64 * WM * WN * KS) __attribute__((amdgpu_waves_per_eu(2))) void k_ldu_sweep(
    const float2* __restrict__ LDU, int64_t lda, const float2* Bin, int64_t ldbin, float2* X, int64_t ldx, int64_t n,
    int64_t nrhs, int nblk, int ct, int* ctr) {
  typedef mg_sg::Tile<NS, BM, BN, WM, WN> Tile;
  __shared__ typename Tile::S sm[KS];
  __shared__ int s_ticket, s_ok;
  const int grp = threadIdx.x / Tile::NT;
  if (threadIdx.x == 0) s_ticket = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int tpr = (kIB / BM) * ct;   // tiles per block row
  const int blk = s_ticket / tpr, w = s_ticket % tpr;
  const int i = FWD ? blk : nblk - 1 - blk;
  int* done = ctr + 1;
  const int64_t i0 = (int64_t)i * kIB + (int64_t)(w / ct) * BM, j0 = (int64_t)(w % ct) * BN;
  const int64_t M = n - i0, N = nrhs - j0;   // valid extents (M <= 0: a row tile past a short last block)
  if (M > 0) {
    Tile tile;
    // group 0 starts from -b_i, the others from 0
    tile.init(Bin + i0 * ldbin + j0, ldbin, M, N, make_float2(-1.f, 0.f), grp == 0);
    bool ok = true;
    const int steps = FWD ? i : nblk - 1 - i;
    auto wait_row = [&](int j) {
      if (threadIdx.x == 0) {
        long long spins = 0;
        int good = 1;
        while (__hip_atomic_load(&done[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < tpr) {
          __builtin_amdgcn_s_sleep(1);
          if (++spins > kSweepSpin) {
            good = 0;
            break;
          }
        }
        s_ok = good;
      }
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      return s_ok != 0;
    };
    const float2* Aj = nullptr;
    const float2* Xj = nullptr;
    int64_t kj = 0, KT = 0;
    auto start_step = [&](int st) {
      const int j = FWD ? st : nblk - 1 - st;
      kj = (n - (int64_t)j * kIB < kIB) ? n - (int64_t)j * kIB : kIB;
      KT = (kj + mg_sg::BK - 1) / mg_sg::BK;
      Aj = LDU + i0 * lda + (int64_t)j * kIB;
      Xj = X + (int64_t)j * kIB * ldx + j0;
      return wait_row(j);
    };
    // One loop over the k-tiles of all steps, KS at a time (group g: k-tile
    // kt + g); the wait for block row j comes before its first fetch.  A
    // group past the block's last k-tile fetches nothing (zeros).
    auto fetch = [&](int64_t kt) {
      tile.template fetch<true>(Aj, lda, Xj, ldx, M, N, kj, (kt + grp) * mg_sg::BK);
    };
    int st = 0;
    int64_t kt = 0;
    if (steps > 0) {
      ok = start_step(0);
      if (ok) fetch(0);
    }
    while (ok && st < steps) {
      __syncthreads();   // previous k-tiles' LDS reads are done
      tile.stash(sm[grp]);
      __syncthreads();
      kt += KS;
      if (kt >= KT) {
        kt = 0;
        if (++st < steps) ok = start_step(st);
      }
      if (ok && st < steps) fetch(kt);
      tile.compute(sm[grp]);
    }
    // group sums into group 0 through LDS, real parts then imaginary parts
    // (each group's own planes hold its partial sums)
    if constexpr (KS > 1) {
      static_assert(sizeof(typename Tile::S) >= sizeof(float) * BM * BN, "the group sums reuse a group's LDS");
#pragma unroll
      for (int part = 0; part < 2; ++part) {
        __syncthreads();
        if (grp > 0) {
          float* buf = reinterpret_cast<float*>(&sm[grp]);
#pragma unroll
          for (int a = 0; a < Tile::TM; ++a)
#pragma unroll
            for (int b = 0; b < Tile::TN; ++b)
#pragma unroll
              for (int r = 0; r < 4; ++r)
                buf[Tile::row_of(a, r) * BN + Tile::col_of(b)] = part ? tile.ci[a][b][r] : tile.cr[a][b][r];
        }
        __syncthreads();
        if (grp == 0) {
#pragma unroll
          for (int g = 1; g < KS; ++g) {
            const float* buf = reinterpret_cast<const float*>(&sm[g]);
#pragma unroll
            for (int a = 0; a < Tile::TM; ++a)
#pragma unroll
              for (int b = 0; b < Tile::TN; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                  const float v = buf[Tile::row_of(a, r) * BN + Tile::col_of(b)];
                  if (part) tile.ci[a][b][r] += v;
                  else tile.cr[a][b][r] += v;
                }
          }
        }
      }
    }
    if (grp == 0) {
      if (!ok) {
        const float nan = __builtin_nanf("");
#pragma unroll
        for (int a = 0; a < Tile::TM; ++a)
#pragma unroll
          for (int b = 0; b < Tile::TN; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) tile.cr[a][b][r] = tile.ci[a][b][r] = nan;
      }
      tile.template store<true>(X + i0 * ldx + j0, ldx, M, N, make_float2(-1.f, 0.f), make_float2(1.f, 0.f),
                                false);
    }
  }
  // every wave's tile stores complete before the barrier after which the
  // row counter is bumped (a workgroup fence emits no such wait on gfx950)
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_fetch_add(&done[i], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sweep launch: counters zeroed by a kernel (no memset node), grid = block
// rows x tiles per block row.
template <int NS, int BM, int BN, int WM, int WN, int KS, bool FWD>
int launch_sweep(int* ctr, const float2* LDU, int64_t lda, const float2* Bin, int64_t ldbin, float2* X, int64_t ldx,
                 int64_t n, int64_t nrhs, hipStream_t s) {
  const int nblk = (int)((n + kIB - 1) / kIB);
  const int ct = (int)((nrhs + BN - 1) / BN);
  const int64_t grid = (int64_t)nblk * (kIB / BM) * ct;
  if (grid > 0x7fffffffLL) return MG_ERR_ARG;
  hipLaunchKernelGGL(k_zero_ints, dim3((unsigned)((nblk + 1 + 255) / 256)), dim3(256), 0, s, ctr, nblk + 1);
  hipLaunchKernelGGL((k_ldu_sweep<NS, BM, BN, WM, WN, KS, FWD>), dim3((unsigned)grid), dim3(64 * WM * WN * KS), 0, s,
                     LDU, lda, Bin, ldbin, X, ldx, n, nrhs, nblk, ct, ctr);
  MG_HIP(hipGetLastError());
  return MG_OK;
}

template <int NS>
int ldu_sweeps(int* ctr, const float2* LDU, int64_t lda, const float2* D, float2* B, int64_t ldb, int64_t nrhs,
               float2* W, int64_t n, hipStream_t s) {
  // BF16x3: 64 x 128 tiles; BF16x6: 64 x 64 (the 64 x 128 tile's LDS would
  // leave one workgroup per CU)
  constexpr int BN = NS == 2 ? 128 : 64;
  constexpr int KS = 1;   // 2 measured slower: at twice the LDS only one workgroup fits per CU
  const int64_t nb = kIB, nblk = (n + nb - 1) / nb, full = n / nb;
  MG_TRY((launch_sweep<NS, 64, BN, 2, 2, KS, true>(ctr, LDU, lda, B, ldb, B, ldb, n, nrhs, s)));
  if (full > 0)
    MG_TRY(mg::cgemm_batched<float>(nb, nrhs, nb, mg::cd{1, 0}, D, nb, nb * nb, B, ldb, nb * ldb, mg::cd{0, 0}, W,
                                    nrhs, nb * nrhs, full, s));
  if (full < nblk) {
    const int64_t r0 = full * nb, h = n - r0;
    MG_TRY(mg::cgemm<float>(h, nrhs, h, mg::cd{1, 0}, D + full * nb * nb, nb, B + r0 * ldb, ldb, mg::cd{0, 0},
                            W + r0 * nrhs, nrhs, s));
  }
  return launch_sweep<NS, 64, BN, 2, 2, KS, false>(ctr, LDU, lda, W, nrhs, B, ldb, n, nrhs, s);
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
int64_t ldu_work_elems(int64_t n) { return 4 * (int64_t)kIB * n + 2 * (int64_t)kIB * kIB; }

int getrf_ldu(Workspace& ws, void* A, int64_t n, int64_t lda, void* dinv, void* work, int* info, hipStream_t s,
              GemmFn gemm, GemmFn update, bool newton) {
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
  auto panel = [&](int64_t k, hipStream_t st) -> int {
    const int64_t r0 = k * nb, h = (n - r0 < nb) ? n - r0 : nb, r1 = r0 + h, r = n - r1;
    float2* Dk = D + k * nb * nb;
    const float2* S = Ap + r0 * lda + r0;
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, st, ws.inv_ctl, 64);
    hipLaunchKernelGGL(k_block_inverse, dim3(kSlabs), dim3(kMwThreads), 0, st, S, lda, (int)h, newton ? X0 : Dk,
                       info, r0, ws.inv_exch, reinterpret_cast<unsigned*>(ws.inv_ctl), ws.inv_ctl + 32);
    MG_HIP(hipGetLastError());
    if (newton) {
      hipLaunchKernelGGL(k_two_eye, dim3((unsigned)((h * nb + 255) / 256)), dim3(256), 0, st, R, (int)h);
      MG_HIP(hipGetLastError());
      MG_TRY(cgemm<float>(h, h, h, cd{-1, 0}, S, lda, X0, nb, cd{1, 0}, R, nb, st));    // R = 2I - S X0
      MG_TRY(cgemm<float>(h, h, h, cd{1, 0}, X0, nb, R, nb, cd{0, 0}, Dk, nb, st));     // D_k = X0 R
    }
    if (r == 0) return MG_OK;
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
  MG_TRY(panel(0, s));
  for (int64_t k = 0; k < nblk; ++k) {
    const int64_t r0 = k * nb, h = (n - r0 < nb) ? n - r0 : nb, r1 = r0 + h, r = n - r1;
    if (r == 0) break;
    const int64_t h2 = (r < nb) ? r : nb;   // next block's order
    // block row and column k+1 first
    MG_TRY(trail(k, r1, h2, r1, r, s));
    MG_TRY(trail(k, r1 + h2, r - h2, r1, h2, s));
    if (ws.s2) {
      MG_HIP(hipEventRecord(ws.evA, s));
      MG_HIP(hipStreamWaitEvent(ws.s2, ws.evA, 0));
      MG_TRY(panel(k + 1, ws.s2));
      MG_HIP(hipEventRecord(ws.evB, ws.s2));
      MG_TRY(trail(k, r1 + h2, r - h2, r1 + h2, r - h2, s));
      MG_HIP(hipStreamWaitEvent(s, ws.evB, 0));
    } else {
      MG_TRY(trail(k, r1 + h2, r - h2, r1 + h2, r - h2, s));
      MG_TRY(panel(k + 1, s));
    }
  }
  return MG_OK;
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
  if (ws && (sweep_ns == 2 || sweep_ns == 3)) {
    if (ws->sweep_cap < nblk + 1) {
      // grows rarely (a larger system than any before on this handle)
      MG_HIP(hipStreamSynchronize(s));
      if (ws->sweep) MG_HIP(hipFree(ws->sweep));
      ws->sweep = nullptr;
      ws->sweep_cap = 0;
      const int64_t cap = ((nblk + 1 + 255) / 256) * 256;
      MG_HIP(hipMalloc(&ws->sweep, sizeof(int) * cap));
      ws->sweep_cap = cap;
    }
    if (sweep_ns == 2) return ldu_sweeps<2>(ws->sweep, Ap, lda, D, Bp, ldb, nrhs, W, n, s);
    return ldu_sweeps<3>(ws->sweep, Ap, lda, D, Bp, ldb, nrhs, W, n, s);
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
