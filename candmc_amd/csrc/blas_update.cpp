// The unmasked bulk of the WY trailing update, A −= Y·Z' (a plain FP64
// GEMM: no triangle, no structure), through the vendor BLAS on the calling
// stream, as an alternative to the hand-written lean kernel.
#include <rocblas/rocblas.h>

#include "cqr_internal.h"

namespace cqr {

int blas_update(cqr_ctx* c, int64_t m, int64_t n, int64_t k, const double* Y, int64_t ldy, const double* Z,
                int64_t ldz, double* A, int64_t lda) {
  if (m <= 0 || n <= 0 || k <= 0) return CQR_OK;
  if (!c->blas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return set_err(c, CQR_ERR_DEVICE, "rocblas_create_handle");
    c->blas = h;
  }
  rocblas_handle h = static_cast<rocblas_handle>(c->blas);
  if (rocblas_set_stream(h, c->stream) != rocblas_status_success) return set_err(c, CQR_ERR_DEVICE, "rocblas_set_stream");
  const double alpha = -1.0, beta = 1.0;
  hipEvent_t ev = prof_begin(c);
  const rocblas_status st = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)m,
                                          (rocblas_int)n, (rocblas_int)k, &alpha, Y, (rocblas_int)ldy, Z,
                                          (rocblas_int)ldz, &beta, A, (rocblas_int)lda);
  prof_end(c, ev, PK_GEMM_UPD, 2.0 * m * n * k, 8.0 * ((double)m * k + (double)k * n + 2.0 * m * n));
  if (st != rocblas_status_success) return set_err(c, CQR_ERR_DEVICE, "rocblas_dgemm: status %d", (int)st);
  return CQR_OK;
}

void blas_destroy(cqr_ctx* c) {
  if (c->blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(c->blas));
  c->blas = nullptr;
}

}  // namespace cqr
