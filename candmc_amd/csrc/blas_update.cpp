// Plain FP64 products of the path (no triangle, no structure) through the
// vendor BLAS on the calling stream: the unmasked bulk of the WY trailing
// update A −= Y·Z' and Y = Q1·U⁻¹ of the Householder reconstruction, as an
// alternative to the library's own MFMA kernels.
#include <rocblas/rocblas.h>

#include "cqr_internal.h"

namespace cqr {

// C (m×n) = alpha·op(A)·B + beta·C, column-major; op(A) = Aᵀ with ta.
int blas_gemm(cqr_ctx* c, int ta, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
              const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int kind) {
  if (m <= 0 || n <= 0 || k <= 0) return CQR_OK;
  if (!c->blas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return set_err(c, CQR_ERR_DEVICE, "rocblas_create_handle");
    c->blas = h;
  }
  rocblas_handle h = static_cast<rocblas_handle>(c->blas);
  if (rocblas_set_stream(h, c->stream) != rocblas_status_success) return set_err(c, CQR_ERR_DEVICE, "rocblas_set_stream");
  hipEvent_t ev = prof_begin(c);
  const rocblas_status st = rocblas_dgemm(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                                          rocblas_operation_none, (rocblas_int)m,
                                          (rocblas_int)n, (rocblas_int)k, &alpha, A, (rocblas_int)lda, B,
                                          (rocblas_int)ldb, &beta, C, (rocblas_int)ldc);
  prof_end(c, ev, kind, 2.0 * m * n * k,
           8.0 * ((double)m * k + (double)k * n + (beta != 0.0 ? 2.0 : 1.0) * m * n));
  if (st != rocblas_status_success) return set_err(c, CQR_ERR_DEVICE, "rocblas_dgemm: status %d", (int)st);
  return CQR_OK;
}

// A (m×n) −= Y·Z (Y m×k, Z k×n).
int blas_update(cqr_ctx* c, int64_t m, int64_t n, int64_t k, const double* Y, int64_t ldy, const double* Z,
                int64_t ldz, double* A, int64_t lda) {
  return blas_gemm(c, 0, m, n, k, -1.0, Y, ldy, Z, ldz, 1.0, A, lda, PK_GEMM_UPD);
}

void blas_destroy(cqr_ctx* c) {
  if (c->blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(c->blas));
  c->blas = nullptr;
}

}  // namespace cqr
