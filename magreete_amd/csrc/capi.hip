// C-ABI: handles, error reporting, factor/solve entry points and element
// conversions.  No C++ exception crosses this boundary; every entry point
// returns an MG_* status.
#include "mg_internal.h"
#include "lu.h"
#include <cstdarg>
#include <cstring>
#include <mutex>

namespace mg {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace mg

struct mg_handle {
  int device = 0;
  mg::Workspace ws;
};

namespace {

__global__ void k_c128_to_c64(const double2* __restrict__ src, int64_t lds, float2* __restrict__ dst, int64_t ldd,
                              int64_t rows, int64_t cols) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  for (int64_t r = blockIdx.y; r < rows; r += gridDim.y) {
    const double2 v = src[r * lds + c];
    dst[r * ldd + c] = make_float2((float)v.x, (float)v.y);
  }
}

__global__ void k_c64_to_c128(const float2* __restrict__ src, int64_t lds, double2* __restrict__ dst, int64_t ldd,
                              int64_t rows, int64_t cols) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  for (int64_t r = blockIdx.y; r < rows; r += gridDim.y) {
    const float2 v = src[r * lds + c];
    dst[r * ldd + c] = make_double2((double)v.x, (double)v.y);
  }
}

}  // namespace

extern "C" {

const char* mg_last_error(void) { return mg::g_err; }

int mg_version(void) { return 1; }

int mg_create(mg_handle** out) {
  if (!out) return MG_ERR_ARG;
  mg_handle* h = new (std::nothrow) mg_handle();
  if (!h) return MG_ERR_NOMEM;
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    mg::set_error("hipGetDevice failed");
    return MG_ERR_HIP;
  }
  const int st = mg::workspace_init(h->ws);
  if (st != MG_OK) {
    mg::workspace_free(h->ws);
    delete h;
    return st;
  }
  *out = h;
  return MG_OK;
}

int mg_destroy(mg_handle* h) {
  if (!h) return MG_OK;
  mg::workspace_free(h->ws);
  delete h;
  return MG_OK;
}

// LU factorisation in place.  prec MG_C128: A is complex double; MG_C64,
// MG_BF16X3 and MG_BF16X6: A is complex float (the BF16 forms run the
// trailing updates on split bf16 MFMA products).  ipiv: n device ints; info: one device int (first zero
// pivot, 1-based, or 0).
int mg_getrf(mg_handle* h, int prec, void* A, int64_t n, int64_t lda, int* ipiv, int* info, void* stream) {
  if (!h || !A || !ipiv || !info) return MG_ERR_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  switch (prec) {
    case MG_C128: return mg::getrf<double>(h->ws, A, n, lda, ipiv, info, s, 256, nullptr);
    case MG_C64: return mg::getrf<float>(h->ws, A, n, lda, ipiv, info, s, 256, nullptr);
    case MG_BF16X3: return mg::getrf<float>(h->ws, A, n, lda, ipiv, info, s, 512, &mg::cgemm_bf16x3_trail);
    case MG_BF16X6: return mg::getrf<float>(h->ws, A, n, lda, ipiv, info, s, 512, &mg::cgemm_bf16x6_trail);
    default: return MG_ERR_ARG;
  }
}

// Solve A X = B in place of B (n x nrhs, row-major) from getrf's factors.
int mg_getrs(mg_handle* h, int prec, const void* LU, int64_t n, int64_t lda, const int* ipiv, void* B, int64_t ldb,
             int64_t nrhs, void* stream) {
  if (!h || !LU || !ipiv || !B) return MG_ERR_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  switch (prec) {
    case MG_C128: return mg::getrs<double>(LU, n, lda, ipiv, B, ldb, nrhs, s);
    case MG_C64:
    case MG_BF16X3:
    case MG_BF16X6: return mg::getrs<float>(LU, n, lda, ipiv, B, ldb, nrhs, s);
    default: return MG_ERR_ARG;
  }
}

// Elements (of the factor's type) of the diagonal-block inverse store for order n.
int64_t mg_dinv_elems(int64_t n) { return n < 0 ? 0 : mg::dinv_elems(n); }

// Inverses of the 256 x 256 diagonal blocks of L and U from getrf's factors.
int mg_getrf_dinv(mg_handle* h, int prec, const void* LU, int64_t n, int64_t lda, void* dinv, void* stream) {
  if (!h || !LU || !dinv || lda < n) return MG_ERR_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  switch (prec) {
    case MG_C128: return mg::diag_inverses<double>(LU, n, lda, dinv, s);
    case MG_C64:
    case MG_BF16X3:
    case MG_BF16X6: return mg::diag_inverses<float>(LU, n, lda, dinv, s);
    default: return MG_ERR_ARG;
  }
}

// Solve with getrf's factors and mg_getrf_dinv's inverses, GEMMs only; work
// holds n x nrhs elements.
int mg_getrs_dinv(mg_handle* h, int prec, const void* LU, int64_t n, int64_t lda, const int* ipiv, const void* dinv,
                  void* B, int64_t ldb, int64_t nrhs, void* work, void* stream) {
  if (!h || !LU || !dinv || !B || !work) return MG_ERR_ARG;   // ipiv NULL: unpivoted factors
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  switch (prec) {
    case MG_C128: return mg::getrs_dinv<double>(LU, n, lda, ipiv, dinv, B, ldb, nrhs, work, s);
    case MG_C64: return mg::getrs_dinv<float>(LU, n, lda, ipiv, dinv, B, ldb, nrhs, work, s);
    case MG_BF16X3: return mg::getrs_dinv<float>(LU, n, lda, ipiv, dinv, B, ldb, nrhs, work, s, &mg::cgemm_bf16x3);
    case MG_BF16X6: return mg::getrs_dinv<float>(LU, n, lda, ipiv, dinv, B, ldb, nrhs, work, s, &mg::cgemm_bf16x6);
    default: return MG_ERR_ARG;
  }
}

int64_t mg_inverse_diag_work_elems(int64_t n) { return n < 0 ? 0 : mg::inverse_diag_work_elems(n); }

// Diagonal blocks of M^{-1} from the pivoted complex128 LU (consumed).
int mg_inverse_diag_blocks(mg_handle* h, int prec, void* LU, int64_t n, int64_t lda, const void* dinv,
                           const int* colperm, int d, const int64_t* sites, int64_t nsites, void* out, void* work,
                           int64_t work_elems, void* stream) {
  if (!h || !LU || !dinv || !colperm || (nsites > 0 && (!sites || !out)) || (n > 0 && !work)) return MG_ERR_ARG;
  if (prec != MG_C128) {
    mg::set_error("inverse diagonal blocks: complex128 factors only (the LDOS needs them to FP64 accuracy)");
    return MG_ERR_UNSUPPORTED;
  }
  return mg::inverse_diag_blocks(h->ws, LU, n, lda, dinv, colperm, d, sites, nsites, out, work, work_elems,
                                 reinterpret_cast<hipStream_t>(stream));
}

int64_t mg_ldu_dinv_elems(int64_t n) { return n < 0 ? 0 : mg::ldu_dinv_elems(n); }
int64_t mg_ldu_work_elems(int64_t n) { return n < 0 ? 0 : mg::ldu_work_elems(n); }

namespace {
// bf16 parts per float of a split precision (0: FP32 products); -1: not an LDU precision
int split_parts(int prec) {
  switch (prec) {
    case MG_C64: return 0;
    case MG_BF16X3: return 2;
    case MG_BF16X6: return 3;
    default: return -1;
  }
}
}  // namespace

int64_t mg_ldu_factor_elems(int prec, int64_t n) {
  const int ns = split_parts(prec);
  return (n < 0 || ns < 0) ? 0 : mg::ldu_factor_elems(ns, n);
}
int64_t mg_ldu_solve_work_elems(int prec, int64_t n, int64_t nrhs) {
  const int ns = split_parts(prec);
  return (n < 0 || nrhs < 0 || ns < 0) ? 0 : mg::ldu_solve_work_elems(ns, n, nrhs);
}

// Block LDU factorisation without pivoting (complex float; MG_C64: FP32 MFMA
// GEMMs, MG_BF16X3 / MG_BF16X6: split-bf16 GEMMs).  The FP32-level forms
// refine each diagonal-block inverse by one FP32 Newton-Schulz step.
int mg_getrf_ldu(mg_handle* h, int prec, void* A, int64_t n, int64_t lda, void* dinv, void* work, int* info,
                 void* stream) {
  if (!h || !A || !dinv || !work || !info) return MG_ERR_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  switch (prec) {
    case MG_C64: return mg::getrf_ldu(h->ws, A, n, lda, dinv, work, info, s, nullptr, nullptr, true, 0);
    case MG_BF16X3:
      return mg::getrf_ldu(h->ws, A, n, lda, dinv, work, info, s, &mg::cgemm_bf16x3, &mg::cgemm_bf16x3_trail, false,
                           2);
    case MG_BF16X6:
      return mg::getrf_ldu(h->ws, A, n, lda, dinv, work, info, s, &mg::cgemm_bf16x6, &mg::cgemm_bf16x6_trail, true,
                           3);
    default: return MG_ERR_ARG;
  }
}

int mg_getrs_ldu(mg_handle* h, int prec, const void* LDU, int64_t n, int64_t lda, const void* dinv, void* B,
                 int64_t ldb, int64_t nrhs, void* work, void* stream) {
  // the split precisions solve from dinv's D inverses and pre-split panels alone (LDU may be null)
  if (!h || !dinv || !B || !work || (!LDU && prec == MG_C64)) return MG_ERR_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  switch (prec) {
    case MG_C64: return mg::getrs_ldu(&h->ws, LDU, n, lda, dinv, B, ldb, nrhs, work, s, nullptr, 0);
    case MG_BF16X3: return mg::getrs_ldu(&h->ws, LDU, n, lda, dinv, B, ldb, nrhs, work, s, &mg::cgemm_bf16x3, 2);
    case MG_BF16X6: return mg::getrs_ldu(&h->ws, LDU, n, lda, dinv, B, ldb, nrhs, work, s, &mg::cgemm_bf16x6, 3);
    default: return MG_ERR_ARG;
  }
}

// Cap on the workgroups of one LU leaf launch for this handle (g <= 0: all CUs).
int mg_set_leaf_blocks(mg_handle* h, int g) {
  if (!h) return MG_ERR_ARG;
  return mg::workspace_leaf_blocks(h->ws, g);
}

// Waits for the stream and reports a failed grid barrier of earlier factorisations.
int mg_check(mg_handle* h, void* stream) {
  if (!h) return MG_ERR_ARG;
  return mg::workspace_check(h->ws, reinterpret_cast<hipStream_t>(stream));
}

// Turn event timing of the LU trailing-update launches on (1) or off (0).
int mg_profile_enable(mg_handle* h, int on) {
  if (!h) return MG_ERR_ARG;
  h->ws.prof.enabled = on != 0;
  h->ws.prof.used = 0;
  return MG_OK;
}

// Launch count, algorithmic flops and summed milliseconds of the trailing
// updates recorded since the last read; synchronises `stream`.
int mg_profile_read(mg_handle* h, void* stream, long long* launches, double* flops, double* ms) {
  if (!h) return MG_ERR_ARG;
  return mg::profile_read(h->ws, reinterpret_cast<hipStream_t>(stream), launches, flops, ms);
}

// As mg_profile_read, per kernel class: element 0 the launches whose
// operands are staged through registers, element 1 those with operands DMA'd
// to LDS (arrays of two).
int mg_profile_read_classes(mg_handle* h, void* stream, long long* launches, double* flops, double* ms) {
  if (!h || !launches || !flops || !ms) return MG_ERR_ARG;
  return mg::profile_read(h->ws, reinterpret_cast<hipStream_t>(stream), launches, flops, ms, true);
}

// dst = (precision of dst) src for a rows x cols block; dir 0: c128 -> c64, 1: c64 -> c128.
int mg_convert(int dir, const void* src, int64_t lds, void* dst, int64_t ldd, int64_t rows, int64_t cols,
               void* stream) {
  if (rows < 0 || cols < 0) return MG_ERR_ARG;
  if (rows == 0 || cols == 0) return MG_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)((cols + 255) / 256), (unsigned)(rows < 65535 ? rows : 65535));
  if (dir == 0)
    hipLaunchKernelGGL(k_c128_to_c64, grid, dim3(256), 0, s, reinterpret_cast<const double2*>(src), lds,
                       reinterpret_cast<float2*>(dst), ldd, rows, cols);
  else if (dir == 1)
    hipLaunchKernelGGL(k_c64_to_c128, grid, dim3(256), 0, s, reinterpret_cast<const float2*>(src), lds,
                       reinterpret_cast<double2*>(dst), ldd, rows, cols);
  else
    return MG_ERR_ARG;
  MG_HIP(hipGetLastError());
  return MG_OK;
}

}  // extern "C"
