// Complex-float GEMM with each real product split into three bf16 MFMA
// products (BF16x3).  Placeholder until the kernel lands.
#include "mg_internal.h"

namespace mg {

int cgemm_bf16x3(int64_t M, int64_t N, int64_t K, cd alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                 cd beta, void* C, int64_t ldc, hipStream_t s) {
  (void)M; (void)N; (void)K; (void)alpha; (void)A; (void)lda; (void)B; (void)ldb; (void)beta; (void)C; (void)ldc;
  (void)s;
  set_error("bf16x3 GEMM not built");
  return MG_ERR_UNSUPPORTED;
}

}  // namespace mg
