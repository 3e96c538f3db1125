// Internal C++ API shared by the translation units of libmagreete_hip.so.
#pragma once
#include "mg_common.h"
#include <hip/hip_runtime.h>
#include <cstdio>
#include <complex>

#define MG_HIP(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      mg::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,                \
                    hipGetErrorString(e_));                                   \
      return MG_ERR_HIP;                                                      \
    }                                                                         \
  } while (0)

#define MG_TRY(expr)                     \
  do {                                   \
    int s_ = (expr);                     \
    if (s_ != MG_OK) return s_;          \
  } while (0)

namespace mg {

void set_error(const char* fmt, ...);

struct cd { double re, im; };

// C = alpha A B + beta C for row-major complex matrices of element T2.
template <typename T>
int cgemm(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
          cd beta, void* C, int64_t ldc, hipStream_t s);

// Batched form: problem b uses A + b sA, B + b sB, C + b sC (element strides).
template <typename T>
int cgemm_batched(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, int64_t sA, const void* B,
                  int64_t ldb, int64_t sB, cd beta, void* C, int64_t ldc, int64_t sC, int64_t batch,
                  hipStream_t s);

// Workspace (elements of T2) with which cgemm_ws splits K for a tall-skinny
// product whose tile grid would leave CUs idle; 0 when it does not split.
template <typename T>
int64_t cgemm_work_elems(int64_t M, int64_t N, int64_t K);
// cgemm with that split: partial products into work, then a fixed-order
// reduction into C.  Without enough workspace it is cgemm.
template <typename T>
int cgemm_ws(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
             cd beta, void* C, int64_t ldc, void* work, int64_t work_elems, hipStream_t s);

// The same GEMM launched under a separate kernel symbol (LU trailing update).
template <typename T>
int cgemm_trail(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                cd beta, void* C, int64_t ldc, hipStream_t s);

// C = alpha A B + beta C with A and B complex float, products as three
// bf16 MFMA products each (hi*hi + hi*lo + lo*hi).
int cgemm_bf16x3(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B,
                 int64_t ldb, cd beta, void* C, int64_t ldc, hipStream_t s);
// The same GEMM under its own kernel symbol (factorisation trailing updates).
int cgemm_bf16x3_trail(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B,
                       int64_t ldb, cd beta, void* C, int64_t ldc, hipStream_t s);
// BF16x6: three-part bf16 splits and the six products with i + j <= 2 per
// real product (FP32-level accuracy).
int cgemm_bf16x6(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B,
                 int64_t ldb, cd beta, void* C, int64_t ldc, hipStream_t s);
int cgemm_bf16x6_trail(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B,
                       int64_t ldb, cd beta, void* C, int64_t ldc, hipStream_t s);

// Compute units of the current device (cached per device).
int cu_count();

}  // namespace mg
