/*
 * CANDMC_QR.h — drop-in C++ declarations of CANDMC's QR hot path, served by
 * libcandmc_qr.so on one MI355X.
 *
 * A caller of the reference that includes "CANDMC.h" for the QR functions
 * includes this header instead and links -lcandmc_qr.  The signatures are
 * those of the reference headers (paths relative to the CANDMC tree):
 *   alg/QR/tsqr/bitree_tsqr.h:18-138, alg/QR/hh_recon/hh_recon.h:13-75,
 *   alg/QR/qr_2d/qr_2d.h:25-188,237-249,293-302, alg/shared/util.h (tau_recon and
 *   the packed-triangle helpers), alg/shared/comm.h:32-84 (CommData_t,
 *   pview).
 * Semantics: matrices are column-major doubles with explicit leading
 * dimensions, as in the reference.  Pointers may be host or device memory
 * (host arrays are staged through the GPU and copied back).  The process
 * grid collapses onto one GPU: a communicator must have np ≤ 1; numPes in
 * the TSQR calls then names the number of virtual row blocks of the tree.
 * Errors abort with a message, as the reference's asserts do.
 */
#ifndef CANDMC_QR_DROPIN_H
#define CANDMC_QR_DROPIN_H

#include <stdint.h>

#ifndef TAU_BLK
#define TAU_BLK 16
#endif

/* alg/shared/comm.h:32-63 without MPI: cm is an opaque handle. */
typedef class CommData {
 public:
  void* cm;
  int np;
  int rank;
  int color;
  int alive;
} CommData_t;

/* alg/shared/comm.h:66-84 */
class pview {
 public:
  int rrow;
  int rcol;
  CommData_t crow;
  CommData_t ccol;
  CommData_t cdiag;
  CommData_t cworld;
};

/* One process-wide GPU context; returns the C-ABI handle (candmc_qr.h). */
struct cqr_ctx;
cqr_ctx* candmc_qr_context();

/* alg/QR/tsqr/bitree_tsqr.h */
void tree_tsqr(double* A, double* buf, int64_t const b, double* tau = 0);
void local_tsqr(double* A, double* R, int64_t const m, int64_t const b, int64_t const lda_A,
                double* tau = 0);
void bitree_tsqr(double* A, int64_t const lda_A, double* R, double* tau, int64_t const m,
                 int64_t const b, int64_t const myRank, int64_t const numPes, int64_t const root,
                 int64_t const req_id, CommData_t cdt, int64_t const output_Y = 0,
                 double* tree_data = 0);
void construct_Q1(double const* Y, int64_t const lda_Y, double const* tau, double* Q1,
                  int64_t const lda_Q, int64_t const m, int64_t const b, int64_t const k,
                  int64_t const myRank, int64_t const numPes, int64_t const root, CommData_t cdt,
                  double* tree_data = 0, double* ID = 0);
void apply_tsqr_QT(double const* Y, int64_t const lda_Y, double const* tau, double* B,
                   int64_t const lda_B, int64_t const m, int64_t const b, int64_t const k,
                   int64_t const myRank, int64_t const numPes, int64_t const root, CommData_t cdt,
                   double* tree_data = 0);

/* alg/QR/hh_recon/hh_recon.h */
void compute_invT_from_W(double const* W, int64_t b, double* invT);
void signed_NLU(double* A, double* R, int64_t b, int64_t lda_A, int64_t lda_R, int64_t* signs);
void recursive_NLU(double* A, double* R, int64_t b, int64_t lda_A, int64_t lda_R, int64_t* signs);
void hh_recon_qr(double* A, int64_t lda_A, int64_t m, int64_t b, double* W, int64_t myRank,
                 int64_t numPes, int64_t root, int64_t req_id, CommData_t cdt);

/* alg/QR/qr_2d/qr_2d.h */
void compute_invT_from_Y(double const* Y, int64_t lda_Y, double* T, int64_t lda_T, int64_t mb,
                         int64_t b, pview* pv);
void compute_invT_from_Y(double const* Y1, int64_t lda_Y1, double const* Y2, int64_t lda_Y2, double* T,
                         int64_t lda_T, int64_t mb, int64_t b1, int64_t b2, pview* pv);
void update_A(double const* Y, int64_t lda_Y, double* A, int64_t lda_A, int64_t m, int64_t k,
              int64_t b, double const* W, pview* pv, double* aggreg_Y, int64_t lda_aY,
              bool W_is_T = false);
void upd_A(double const* Ybuf, int64_t lda_Y, double* A, int64_t lda_A, int64_t mb, int64_t kb,
           int64_t b, double const* W, pview* pv, bool W_is_T = false);
void QR_2D(double* A, int64_t lda_A, int64_t m, int64_t k, int64_t b, pview* pv, double* aggreg_Y,
           int64_t lda_aY);
void QR_2D_2D(double* A, int64_t lda_A, int64_t m, int64_t k, int64_t b, int64_t b_sub, pview* pv,
              double* aggreg_Y, int64_t lda_aY, int* desc_A = 0, double* org_A = 0, int64_t IA = 1,
              int64_t JA = 1);
void QR_tree_2D(double* A, int64_t const lda_A, int64_t const m, int64_t const k, int64_t const b,
                int64_t const myRank, int64_t const numPes, int64_t const root_row,
                int64_t const root_col, CommData_t cdt_row, CommData_t cdt_col, CommData_t cdt_world,
                int64_t const _stop_at = -1);
void QR_2D_pipe(double* A, int64_t lda_A, int64_t m, int64_t k, int64_t b, pview* pv,
                double* last_Y = 0, int64_t lda_lY = 0, double* last_W = 0, double* my_last_W = 0);

/* alg/shared/util.h */
void tau_recon(const char structure, const int ncol, const int nrow, const int lda, double const* Y,
               double* tau);
void pack_upper(double const* R_in, double* R_out, int const b, int const lda_in);
void unpack_upper(double const* R_in, double* R_out, int const b, int const lda_out);
void pack_lower(double const* R_in, double* R_out, int const b, int const lda_in, int const has_diag);
void unpack_lower(double const* R_in, double* R_out, int const b, int const lda_out,
                  int const has_diag);

#endif
