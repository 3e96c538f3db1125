// LU factorisation and solve drivers (internal).
#pragma once
#include "mg_internal.h"
#include <vector>

namespace mg {

constexpr int kLeaf = 32;              // widest leaf of the triangular solves
// widest panel factored by one cooperative leaf launch; leaves of 64 to 256
// columns (fewer swap / triangular-solve / GEMM launches in the panel) were
// measured on the 12288-row factorisation: the same 138 ms factor, and the
// longer leaf launches slowed the trailing update they share CUs with
// (solve 160 ms against 158 ms)
constexpr int kGetf2MaxW = 32;
constexpr int kMaxGetf2Blocks = 256;   // workgroups of one leaf launch (<= one per CU)
// aim for at least this many rows per workgroup (MG_GETF2_ROWS overrides it);
// the 12288-row factorisation: 141.5 ms at 32 (256 workgroups), 138.9 at 48,
// 137.0-137.4 at 64, 136.7-137.1 at 96 (128 workgroups): flat above 64
constexpr int kGetf2MinRows = 64;
// dynamic LDS per leaf workgroup: small enough that a leaf workgroup shares its
// CU with a trailing-update workgroup (the look-ahead runs both at once);
// measured on the 12288-row factorisation: 138 ms at 48 KB, 158 ms at 96 KB,
// 161 ms at 144 KB
constexpr int kGetf2Lds = 48 * 1024;

// Cross-workgroup exchange area of the leaf kernel (double-buffered by
// column parity).  The barrier counters come in two sets used by alternate
// leaf launches of a handle: a launch counts in one set and clears the
// other for the next launch, so no zeroing launch precedes a leaf.  Each set
// holds its launch's failure word; a failure also sets the sticky flag,
// which only the handle's check reads and clears.
struct Getf2Work {
  double* cand_val = nullptr;   // [2][G] best |re|+|im| per workgroup
  int* cand_row = nullptr;      // [2][G] its row
  void* cand_data = nullptr;    // [2][G][kGetf2MaxW] the candidate rows
  void* rowj = nullptr;         // [2][kGetf2MaxW] row j before the swap
  unsigned* bar = nullptr;      // the set this launch counts in (kernel argument); in the handle: both sets
  unsigned* bar_next = nullptr; // the set this launch clears (kernel argument)
  int* err = nullptr;           // sticky barrier-failure flag
};

// Optional timing of the LU trailing-update launches with HIP events on the
// launching stream: algorithmic flops (8 m n k per complex GEMM) and elapsed
// time per launch.
struct UpdateProfile {
  bool enabled = false;
  std::vector<hipEvent_t> ev;   // start/stop pairs, reused
  std::vector<double> flops;    // per recorded launch
  std::vector<int> cls;         // per recorded launch: 1 DMA-staged operands, 0 register-staged
  size_t used = 0;
};

struct Workspace {
  bool ready = false;
  int cus = 0;
  int coop_blocks = 0;
  Getf2Work getf2;
  int leaf_parity = 0;             // the barrier set the next leaf counts in
  UpdateProfile prof;
  hipStream_t s2 = nullptr;        // look-ahead panel stream (default priority)
  hipStream_t s3 = nullptr;        // a third stream beside the look-ahead (default priority)
  hipStream_t s4 = nullptr;        // split LDU: Uh panels, which only the solves read (default priority)
  hipEvent_t evA = nullptr, evB = nullptr, evC = nullptr, evD = nullptr, evE = nullptr, evF = nullptr,
             evG = nullptr, evH = nullptr;
  float2* inv_exch = nullptr;      // LDU block inverse: two 256 x 32 published slabs
  int* inv_ctl = nullptr;          // LDU block inverse: two sets of barrier counter [0] and failure flag [32]
  int inv_parity = 0;              // the set the next block inverse uses (it clears the other)
  int* sweep = nullptr;            // two sets of ticket + per-block-row counters of the LDU solve sweeps
  int64_t sweep_cap = 0;           // ints per set
  int sweep_parity = 0;            // the set the next sweep uses (it clears the other)
};

typedef int (*GemmFn)(int64_t, int64_t, int64_t, cd, const void*, int64_t, const void*, int64_t, cd, void*, int64_t,
                      hipStream_t);

// fn(...) on s, timed with HIP events on s when ws.prof is enabled
// (algorithmic flops 8 M N K per launch).
int timed_update(Workspace& ws, GemmFn fn, int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda,
                 const void* B, int64_t ldb, cd beta, void* C, int64_t ldc, hipStream_t s);
// launch() on s, timed with HIP events on s when ws.prof is enabled (flops:
// the launch's algorithmic count).
// cls: the kernel class recorded with the launch (1: operands DMA'd to LDS,
// 0: staged through registers).
template <typename F>
int timed_launch(Workspace& ws, double flops, hipStream_t s, F&& launch, int cls = 0) {
  UpdateProfile& pr = ws.prof;
  if (pr.enabled) {
    if (pr.used * 2 + 2 > pr.ev.size()) {
      hipEvent_t a, b;
      MG_HIP(hipEventCreate(&a));
      MG_HIP(hipEventCreate(&b));
      pr.ev.push_back(a);
      pr.ev.push_back(b);
    }
    MG_HIP(hipEventRecord(pr.ev[pr.used * 2], s));
  }
  MG_TRY(launch());
  if (pr.enabled) {
    MG_HIP(hipEventRecord(pr.ev[pr.used * 2 + 1], s));
    if (pr.flops.size() <= pr.used) pr.flops.resize(pr.used + 1);
    if (pr.cls.size() <= pr.used) pr.cls.resize(pr.used + 1);
    pr.flops[pr.used] = flops;
    pr.cls[pr.used] = cls;
    ++pr.used;
  }
  return MG_OK;
}
// Sum of recorded update launches since the last read (synchronises the
// stream); with per_class, arrays of two: class 0 (register-staged) and
// class 1 (DMA-staged) separately.
int profile_read(Workspace& ws, hipStream_t s, long long* launches, double* flops, double* ms,
                 bool per_class = false);
// Whether the FP64 GEMM of this shape runs the DMA-staged kernel (gemm.hip).
bool zgemm_uses_dma(int64_t M, int64_t N, int64_t K, int64_t batch = 1);

int workspace_init(Workspace& ws);
void workspace_free(Workspace& ws);
// Cap on the workgroups of one LU leaf launch (g <= 0: the CU count).  Solves
// that run concurrently on one device keep the sum of their caps within the
// CU count, so that every leaf can become resident.
int workspace_leaf_blocks(Workspace& ws, int g);
// Synchronises the stream, reports a failed grid barrier of the LU leaf
// since the last check and clears that report.
int workspace_check(Workspace& ws, hipStream_t s);

template <typename T>
int laswp(void* A, int64_t lda, int64_t c0, int64_t c1, int64_t k1, int64_t k2, const int* ipiv, hipStream_t s);
template <typename T>
int trsm_lower_unit(const void* L, int64_t ldl, int64_t nb, void* B, int64_t ldb, int64_t ncols, hipStream_t s);
template <typename T>
int trsm_upper(const void* U, int64_t ldu, int64_t nb, void* B, int64_t ldb, int64_t ncols, hipStream_t s);
// update: GEMM used for the trailing update (nullptr: the FP64/FP32 MFMA GEMM of T).
template <typename T>
int getrf(Workspace& ws, void* A, int64_t n, int64_t lda, int* ipiv, int* info, hipStream_t s, int nb,
          GemmFn update);
template <typename T>
int getrs(const void* LU, int64_t n, int64_t lda, const int* ipiv, void* B, int64_t ldb, int64_t nrhs,
          hipStream_t s);
template <typename T>
int trsm_lower_unit_b(const void* L, int64_t ldl, int64_t sL, int64_t nb, void* B, int64_t ldb, int64_t sB,
                      int64_t ncols, int64_t batch, hipStream_t s);

// Block size of the stored diagonal-block inverses used by getrs_dinv.
constexpr int64_t kInvBlock = 256;
inline int64_t dinv_elems(int64_t n) { return 2 * ((n + kInvBlock - 1) / kInvBlock) * kInvBlock * kInvBlock; }
template <typename T>
int diag_inverses(const void* LU, int64_t n, int64_t lda, void* dinv, hipStream_t s);
// Block LDU without pivoting (complex float): inverses of the 256 x 256
// diagonal blocks of the Schur complements in dinv (ldu_dinv_elems), scaled
// off-diagonal blocks in place, work: ldu_work_elems(n) elements; the solve
// needs n x nrhs work elements.  gemm: GEMM of the panels and block solves,
// update: GEMM of the trailing updates (nullptr: the FP32 MFMA GEMMs);
// newton: refine each block inverse by one FP32 Newton-Schulz step.
// split_ns (2: BF16x3, 3: BF16x6; 0 for FP32 products): the trailing
// updates and the solves read the panels pre-split, which the factorisation
// stores after the D inverses (dinv then holds ldu_factor_elems(split_ns, n)
// elements) and the solve's work holds ldu_solve_work_elems(split_ns, n, nrhs).
int64_t ldu_dinv_elems(int64_t n);
int64_t ldu_factor_elems(int split_ns, int64_t n);
int64_t ldu_work_elems(int64_t n);
int64_t ldu_solve_work_elems(int split_ns, int64_t n, int64_t nrhs);
int getrf_ldu(Workspace& ws, void* A, int64_t n, int64_t lda, void* dinv, void* work, int* info, hipStream_t s,
              GemmFn gemm, GemmFn update, bool newton, int split_ns);
// sweep_ns: 2 / 3 run the forward and backward solves each as one dataflow
// launch with BF16x3 / BF16x6 products over the pre-split panels (counters
// in ws); 0 runs one GEMM launch per block column (gemm, nullptr: the FP32
// MFMA GEMM).
int getrs_ldu(Workspace* ws, const void* LDU, int64_t n, int64_t lda, const void* dinv, void* B, int64_t ldb,
              int64_t nrhs, void* work, hipStream_t s, GemmFn gemm, int sweep_ns);

// d x d diagonal blocks of M^{-1} for the listed sites (block s of rows and
// columns d sites[s] .. d sites[s] + d - 1), from the complex128 pivoted LU
// (getrf) and its diagonal-block inverses (diag_inverses), which are
// overwritten by U^{-1} and L^{-1}.  colperm[c]: the column of L^{-1} P
// (P M = L U) that equals column c of L^{-1} P's product with the swaps,
// i.e. the row that ipiv's swaps move row c to.  out: nsites d^2 complex128
// (row-major blocks); work: inverse_diag_work_elems(n) complex128.
int64_t inverse_diag_work_elems(int64_t n);
int inverse_diag_blocks(Workspace& ws, void* LU, int64_t n, int64_t lda, const void* dinv, const int* colperm, int d,
                        const int64_t* sites, int64_t nsites, void* out, void* work, int64_t work_elems,
                        hipStream_t s);

// gemm: GEMM of the block solves (nullptr: the FP64/FP32 MFMA GEMM of T).
template <typename T>
int getrs_dinv(const void* LU, int64_t n, int64_t lda, const int* ipiv, const void* dinv, void* B, int64_t ldb,
               int64_t nrhs, void* work, hipStream_t s, GemmFn gemm = nullptr);

}  // namespace mg
