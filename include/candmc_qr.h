/*
 * candmc_qr.h — C-ABI of the MI355X communication-avoiding QR hot path.
 *
 * Every compute entry point takes device pointers to column-major FP64
 * matrices with an explicit leading dimension, is asynchronous on the
 * context's HIP stream, and returns a status code (CQR_OK on success).
 * The MPI process grid of the reference collapses onto one GPU: the
 * binary reduction tree runs over "virtual ranks" (leaf row blocks) in
 * HBM, and every communicator argument disappears.  Each declaration
 * names the reference function it replaces (paths relative to the
 * CANDMC source tree).
 */
#ifndef CANDMC_QR_H
#define CANDMC_QR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cqr_ctx cqr_ctx;

enum cqr_status {
  CQR_OK = 0,
  CQR_ERR_INVALID = 1,     /* bad shape, leading dimension or pointer       */
  CQR_ERR_UNSUPPORTED = 2, /* outside the kernels' limits (e.g. b > 128)    */
  CQR_ERR_NOMEM = 3,       /* workspace could not be allocated              */
  CQR_ERR_DEVICE = 4,      /* a HIP runtime call or kernel launch failed    */
  CQR_ERR_NONFINITE = 5,   /* input holds NaN/Inf (only with check_finite),
                              or a least-squares R / iterate is not finite  */
  CQR_ERR_NOCONVERGE = 6,  /* least-squares refinement ended above tol: X
                              holds the last iterate                       */
  CQR_ERR_CANCELLED = 7    /* a served request dropped before it started
                              (the server was destroyed)                   */
};

/* Arithmetic of the trailing (compact-WY) update; panels are always FP64. */
enum cqr_prec {
  CQR_PREC_F64 = 0,  /* FP64 MFMA                                          */
  CQR_PREC_F32 = 1,  /* FP32 MFMA, FP32 accumulation, FP64 matrix in HBM    */
  CQR_PREC_BF16 = 2  /* BF16 MFMA, FP32 accumulation, FP64 matrix in HBM    */
};

/* Largest panel width the FP64 panel kernels accept. */
#define CQR_MAX_B 128

/* ---- context, stream and workspace (no reference counterpart: the
 *      reference mallocs scratch inside every call, e.g.
 *      alg/QR/hh_recon/hh_recon.cxx:147-151) ---------------------------- */
int cqr_create(cqr_ctx** ctx, int device);
int cqr_destroy(cqr_ctx* ctx);
int cqr_set_stream(cqr_ctx* ctx, void* hip_stream);   /* NULL = default stream */
void* cqr_get_stream(const cqr_ctx* ctx);
/* Waits for the context's stream; returns CQR_ERR_DEVICE if a TSQR tree
 * node launched since the last sync gave up waiting for a child (its
 * results are then invalid). */
int cqr_sync(cqr_ctx* ctx);
int cqr_reserve(cqr_ctx* ctx, size_t bytes);          /* pre-grow the arena   */
/* Waits for the context's streams and frees its workspace arena (it grows
 * again on the next call that needs it).                                 */
int cqr_trim(cqr_ctx* ctx);
size_t cqr_workspace_bytes(const cqr_ctx* ctx);
int cqr_set_check_finite(cqr_ctx* ctx, int on);       /* scan inputs first    */
int cqr_set_leaf_rows(cqr_ctx* ctx, int64_t rows);    /* 0 = automatic        */
/* QR_2D lookahead (on by default): the update of the next panel's columns
 * runs first, then the rest of the trailing update proceeds on a second
 * stream of the context while the next panel is factored (the pipeline of
 * alg/QR/qr_2d/qr_2d.cxx:378-635, QR_2D_pipe).  Results differ from the
 * non-lookahead order only in the summation order of Yᵀ·A.               */
int cqr_set_lookahead(cqr_ctx* ctx, int on);
/* Diagnostic: bounds every wait between workgroups of one launch (TSQR tree,
 * construct_Q1) to `polls` polls; 0 restores the default (2²², far beyond
 * any real wait).  A wait that gives up makes its launch finish with what
 * it has, and the next cqr_sync returns CQR_ERR_DEVICE (and clears the
 * flag); a test uses a tiny bound to exercise that path.  No reference
 * counterpart.                                                          */
int cqr_set_wait_limit(cqr_ctx* ctx, int polls);
/* Kernel timing: while enabled, the library records a HIP event pair on
 * its stream around every kernel it launches; cqr_profile_read waits for
 * the stream and returns, for one kernel class, the summed event time,
 * the algorithmic flops and HBM bytes of those launches and their count.
 * Classes: 0 Yᵀ·A split-K GEMM, 1 A −= Y·Z' update GEMM, 2 small b×b
 * GEMMs on MFMA, 3 split-K reduction, 4 leaf QR, 5 TSQR tree, 6 HR core,
 * 7 batched small GEMMs of the panel.                                    */
int cqr_profile_enable(cqr_ctx* ctx, int on);
int cqr_profile_read(cqr_ctx* ctx, int kind, double* ms, double* flops, double* bytes,
                     int64_t* launches);
int cqr_profile_reset(cqr_ctx* ctx);
const char* cqr_status_string(int status);
const char* cqr_last_error(const cqr_ctx* ctx);

/* ---- host-only planning (no device needed) ----------------------------- */
/* Leaf layout the TSQR uses for an m×b panel split over at least
 * `nleaves_hint` row blocks (0 = automatic): the reference's binary tree
 * (alg/QR/tsqr/bitree_tsqr.cxx:172) over `*nleaves` virtual ranks.      */
int cqr_plan_tsqr(int64_t m, int64_t b, int64_t nleaves_hint, int64_t leaf_rows,
                  int64_t* nleaves, int64_t* max_leaf_rows, int64_t* nlevels);
/* Doubles needed for `tree_data` of cqr_bitree_tsqr (replaces the
 * 2b²(ln P + 2) sizing of alg/QR/hh_recon/hh_recon.cxx:151).            */
int64_t cqr_tree_data_doubles(int64_t m, int64_t b, int64_t nleaves);

/* ---- a1  local_tsqr  (alg/QR/tsqr/bitree_tsqr.cxx:48-71) ---------------
 * dgeqrf of one m×b block: A := [V\R], tau[b].  T (b×b upper, the compact
 * WY factor, nullable) is returned too.  Blocks larger than one CU's leaf
 * capacity are factored by TSQR + Householder reconstruction, which
 * yields the same [V\R] and tau.                                          */
int cqr_local_tsqr(cqr_ctx* ctx, double* A, int64_t lda, int64_t m, int64_t b,
                   double* tau, double* T);

/* ---- a2  tree_tsqr  (alg/QR/tsqr/bitree_tsqr.cxx:20-34) ----------------
 * QR of two stacked b×b upper triangles held in a 2b×b buffer (ld 2b), as
 * dtpqrt(b,b,l=b,nb=min(b,16)): top := R, bottom := V2 (upper), T := the
 * nb×b blocked triangular factor (ld nb, nullable).                       */
int cqr_tree_tsqr(cqr_ctx* ctx, double* A2b, int64_t b, double* T);

/* ---- a3  bitree_tsqr  (alg/QR/tsqr/bitree_tsqr.cxx:91-248) -------------
 * TSQR of the m×b panel over `nleaves` virtual ranks (0 = automatic).
 * R (b×b, zero below the diagonal), tau (nleaves×b: each leaf's dgeqrf
 * tau), A := leaf [V\R] blocks when output_Y != 0, tree_data := opaque
 * per-leaf and per-node factors (size from cqr_tree_data_doubles).        */
int cqr_bitree_tsqr(cqr_ctx* ctx, double* A, int64_t lda, double* R, double* tau,
                    int64_t m, int64_t b, int64_t nleaves, int output_Y,
                    double* tree_data);

/* ---- a4  construct_Q1  (alg/QR/tsqr/construct_tsqr_Q.cxx:34-242) -------
 * Explicit first k = b columns of the TSQR Q from the outputs of
 * cqr_bitree_tsqr (same m, b, nleaves).                                   */
int cqr_construct_Q1(cqr_ctx* ctx, const double* Y, int64_t lda_Y, const double* tau,
                     double* Q1, int64_t lda_Q, int64_t m, int64_t b, int64_t k,
                     int64_t nleaves, const double* tree_data);
/* construct_Q1 in full (alg/QR/tsqr/construct_tsqr_Q.cxx:34-242 with
 * ID and any k): Q1 (m×k, k ≤ m) := Q·[ID; 0] with ID b×k (ld ld_ID), or,
 * with ID == NULL, the first k columns of the TSQR's m×m Q.  Q is the
 * product of the leaves' and tree nodes' reflectors from cqr_bitree_tsqr.
 * (The reference's multi-rank branch seeds identity entries only in each
 * rank's top b rows, so its columns b..k−1 are not Q's for k > b; this
 * follows its one-rank dorgqr branch, :87-102, which returns Q's.)        */
int cqr_construct_Q1_id(cqr_ctx* ctx, const double* Y, int64_t lda_Y, const double* tau,
                        double* Q1, int64_t lda_Q, int64_t m, int64_t b, int64_t k,
                        int64_t nleaves, const double* tree_data, const double* ID,
                        int64_t ld_ID);

/* ---- apply_tsqr_QT  (alg/QR/tsqr/apply_tsqr_QT.cxx:31-165) -------------
 * B (m×k, ld lda_B) := Qᵀ·B with the implicit Q of cqr_bitree_tsqr (Y =
 * its leaf [V\R] blocks, tree_data, same m, b, nleaves).  Row block l of
 * the result is leaf l's rows; for B = the factored panel's input, the top
 * b rows of leaf 0 receive R and every other row is zero.  tau is unused
 * (tree_data holds the leaves' T).  cqr_apply_tsqr_Q applies Q instead.  */
int cqr_apply_tsqr_QT(cqr_ctx* ctx, const double* Y, int64_t lda_Y, const double* tau,
                      double* B, int64_t lda_B, int64_t m, int64_t b, int64_t k,
                      int64_t nleaves, const double* tree_data);
int cqr_apply_tsqr_Q(cqr_ctx* ctx, const double* Y, int64_t lda_Y, const double* tau,
                     double* B, int64_t lda_B, int64_t m, int64_t b, int64_t k,
                     int64_t nleaves, const double* tree_data);

/* ---- QR_tree_2D  (alg/QR/qr_2d/qr_tree_2d.cxx:30-124) ------------------
 * Tree-updated QR of the m×k matrix A over the first stop_at columns
 * (stop_at < 0: min(m, k)), panels of b: each panel is factored by
 * bitree TSQR over `nleaves` virtual row blocks (0 = automatic), its R is
 * written over the top of the panel, and the trailing columns receive
 * Qᵀ by cqr_apply_tsqr_QT; the next panel starts b rows and b columns
 * further.  A := per-panel leaf [V\R] blocks with each panel's R on top
 * and the transformed trailing columns; the per-panel tree factors are
 * not kept (as in the reference).                                          */
int cqr_qr_tree_2d(cqr_ctx* ctx, double* A, int64_t lda, int64_t m, int64_t k, int64_t b,
                   int64_t stop_at, int64_t nleaves);

/* ---- tree_data in the reference's layout --------------------------------
 * Virtual rank `rank`'s share of tree_data as bitree_tsqr.cxx:213-217
 * stores it: for every level at which the rank receives, the dtpqrt T
 * (min(b,16)×b, ld min(b,16)) followed by the packed upper V2
 * (b(b+1)/2 doubles, alg/shared/util.cxx pack_upper).  *ndoubles (nullable)
 * receives the size; out == NULL only queries it.                          */
int cqr_tree_data_export(cqr_ctx* ctx, const double* tree_data, int64_t m, int64_t b,
                         int64_t nleaves, int64_t rank, double* out, int64_t* ndoubles);

/* ---- a5  signed_NLU / recursive_NLU  (alg/QR/hh_recon/hh_recon.cxx:43-111)
 * LU without pivoting of the b×b block A − S, S = diag(s) chosen on the fly
 * (s_i = −1 if the pivot is positive, else +1); R's rows are scaled by S.
 * signs (device, int64) may be NULL.                                      */
int cqr_signed_nlu(cqr_ctx* ctx, double* A, int64_t lda, double* R, int64_t lda_R,
                   int64_t b, int64_t* signs);
int cqr_recursive_nlu(cqr_ctx* ctx, double* A, int64_t lda, double* R, int64_t lda_R,
                      int64_t b, int64_t* signs);

/* ---- a6  hh_recon_qr  (alg/QR/hh_recon/hh_recon.cxx:129-321) -----------
 * TSQR + Householder reconstruction of the m×b panel: A := [Y\R] (Y unit
 * lower, implicit diagonal), W := U·S = −T·Y1ᵀ (b×b upper).  The _ex form
 * also returns T (b×b upper, Q = I − Y·T·Yᵀ) and T⁻ᵀ (b×b lower, the form
 * upd_A consumes); either may be NULL.                                    */
int cqr_hh_recon_qr(cqr_ctx* ctx, double* A, int64_t lda, int64_t m, int64_t b,
                    double* W, int64_t nleaves);
int cqr_hh_recon_qr_ex(cqr_ctx* ctx, double* A, int64_t lda, int64_t m, int64_t b,
                       double* W, double* T, double* invT_T, int64_t nleaves);

/* ---- hh_recon_qr without allocation (no counterpart in the reference; for
 *      callers that cannot allocate device memory: code under HIP graph
 *      capture, serving loops, harnesses that forbid hipMalloc) ------------
 * cqr_hh_recon_qr_ex on a fixed scratch buffer of
 * cqr_noalloc_scratch_bytes() bytes (96 MiB) that each GPU holds from
 * library load to process end; panels whose workspace exceeds it (at
 * b = 128 from m = 24 321 rows, at b = 64 from m = 111 947) return
 * CQR_ERR_UNSUPPORTED and leave A untouched.  The
 * call allocates nothing and does not scan the input (check_finite is not
 * applied).  All calls on one GPU share the buffer: calls that overlap in
 * time on one GPU race.                                                   */
size_t cqr_noalloc_scratch_bytes(void);
int cqr_hh_recon_qr_noalloc(cqr_ctx* ctx, double* A, int64_t lda, int64_t m, int64_t b,
                            double* W, double* T, double* invT_T, int64_t nleaves);

/* ---- a7  compute_invT_from_W  (alg/QR/hh_recon/hh_recon.cxx:26-31) -----
 * invT holds Y1 (b×b, ld b) on entry and T⁻ᵀ = −W⁻ᵀ·Y1 on exit.           */
int cqr_compute_invT_from_W(cqr_ctx* ctx, const double* W, int64_t b, double* invT);

/* ---- a8  compute_invT_from_Y  (alg/QR/qr_2d/qr_2d.cxx:22-60) ------------
 * T (b×b, ld lda_T) := tril(YᵀY) with the diagonal halved, Y explicit mb×b. */
int cqr_compute_invT_from_Y(cqr_ctx* ctx, const double* Y, int64_t lda_Y, double* T,
                            int64_t lda_T, int64_t mb, int64_t b);

/* ---- a8  compute_invT_from_Y, two-Y form  (alg/QR/qr_2d/qr_2d.cxx:76-106,
 *      decl alg/QR/qr_2d/qr_2d.h:47) ----------------------------------------
 * T (b1×b2, ld lda_T) := Y1ᵀ·Y2 with Y1 mb×b1 (ld lda_Y1) and Y2 mb×b2
 * (ld lda_Y2): the off-diagonal block of an aggregated T⁻ᵀ.  mb = 0 gives
 * a zero T.                                                                */
int cqr_compute_invT_from_Y2(cqr_ctx* ctx, const double* Y1, int64_t lda_Y1, const double* Y2,
                             int64_t lda_Y2, double* T, int64_t lda_T, int64_t mb, int64_t b1,
                             int64_t b2);

/* ---- a9  upd_A  (alg/QR/qr_2d/qr_2d.cxx:224-282) ------------------------
 * A (mb×kb) := A − Y·(T⁻ᵀ)⁻¹·(Yᵀ·A) with Y explicit (mb×b).  W selects T⁻ᵀ
 * as in the reference: NULL → compute_invT_from_Y(Y); W_is_T → W is T⁻ᵀ;
 * otherwise W is the hh_recon W and T⁻ᵀ = compute_invT_from_W(W, Y1).     */
int cqr_upd_A(cqr_ctx* ctx, const double* Y, int64_t lda_Y, double* A, int64_t lda_A,
              int64_t mb, int64_t kb, int64_t b, const double* W, int W_is_T);

/* ---- a10 update_A  (alg/QR/qr_2d/qr_2d.cxx:124-177) --------------------
 * Builds the unit-lower Y from the panel [Y\R] (m×b, ld lda_Y) and applies
 * upd_A to the m×k block A; copies Y to aggreg_Y (ld lda_aY) if non-NULL.  */
int cqr_update_A(cqr_ctx* ctx, const double* Y, int64_t lda_Y, double* A, int64_t lda_A,
                 int64_t m, int64_t k, int64_t b, const double* W,
                 double* aggreg_Y, int64_t lda_aY, int W_is_T);

/* ---- a11 QR_2D  (alg/QR/qr_2d/qr_2d.cxx:299-363) ------------------------
 * Right-looking blocked QR with TSQR+HR panels of width b: A (m×k, m ≥ k)
 * := [Y\R] in dgeqrf layout (no tau; cqr_tau_recon rebuilds it).          */
int cqr_qr_2d(cqr_ctx* ctx, double* A, int64_t lda, int64_t m, int64_t k, int64_t b,
              double* aggreg_Y, int64_t lda_aY);

/* ---- a12 QR_2D_2D  (alg/QR/qr_2d/qr_2d.cxx:772-897) ---------------------
 * Two-level blocking: QR_2D with b_sub panels on b-wide blocks, then one
 * aggregated rank-b update with T⁻ᵀ from compute_invT_from_Y.             */
int cqr_qr_2d_2d(cqr_ctx* ctx, double* A, int64_t lda, int64_t m, int64_t k, int64_t b,
                 int64_t b_sub, double* aggreg_Y, int64_t lda_aY);

/* ---- a14 tau_recon and packed triangles (alg/shared/util.cxx:263-465) --- */
int cqr_tau_recon(cqr_ctx* ctx, char structure, int64_t ncol, int64_t nrow, int64_t lda,
                  const double* Y, double* tau);
int cqr_pack_upper(cqr_ctx* ctx, const double* R, double* P, int64_t b, int64_t lda);
int cqr_unpack_upper(cqr_ctx* ctx, const double* P, double* R, int64_t b, int64_t lda);
int cqr_pack_lower(cqr_ctx* ctx, const double* T, double* P, int64_t b, int64_t lda,
                   int has_diag);
int cqr_unpack_lower(cqr_ctx* ctx, const double* P, double* T, int64_t b, int64_t lda,
                     int has_diag);

/* ---- f1 least-squares path (no reference counterpart except the
 *      implicit Qᵀ application of alg/QR/tsqr/apply_tsqr_QT.cxx:31-165) --- */
/* QR_2D of an m×k matrix (m ≥ k) that also returns the panel factors:
 * Tt (⌈k/b⌉ blocks of b·b doubles; block p holds Tᵀ of panel p with ld equal
 * to that panel's width), and runs the trailing update in `prec`
 * (CQR_PREC_F64, CQR_PREC_F32 or CQR_PREC_BF16; panels are always FP64). */
int cqr_geqrf2d(cqr_ctx* ctx, double* A, int64_t lda, int64_t m, int64_t k, int64_t b,
                double* Tt, int prec);
/* B (m×nrhs, ld ldb) := Qᵀ·B (trans = 1) or Q·B (trans = 0), with Q from
 * cqr_geqrf2d ([Y\R] in YR, its Tt, the same m, k, b).                  */
int cqr_apply_qt(cqr_ctx* ctx, const double* YR, int64_t ldy, const double* Tt, int64_t m,
                 int64_t k, int64_t b, double* B, int64_t ldb, int64_t nrhs, int trans);
/* B(0:n, :) := R⁻¹·B (trans = 0) or R⁻ᵀ·B (trans = 1), R the upper
 * triangle of the leading n×n block of YR.                               */
int cqr_trsm_r(cqr_ctx* ctx, const double* YR, int64_t ldy, int64_t n, double* B, int64_t ldb,
               int64_t nrhs, int trans);
/* min ‖A·X − B‖ for m ≥ n, 1 ≤ nrhs ≤ 16: QR of a copy of A (A itself is
 * not modified) with the trailing update in `prec`, then up to max_iter
 * steps of iterative refinement with FP64 residuals: corrected
 * semi-normal-equation steps (dx = R⁻¹R⁻ᵀAᵀr) while each shrinks the
 * update at least 8×, then the augmented system; stopping when every
 * column's update is ≤ tol relative to the solution, or when the
 * augmented steps stop shrinking (by less than 10% per step).
 * Returns CQR_OK only when the last relative update is ≤ tol (or
 * max_iter = 0); CQR_ERR_NOCONVERGE when the refinement ended above tol
 * (X holds the last iterate); CQR_ERR_NONFINITE when R or an iterate is
 * not finite.  With a mixed `prec` and escalation on (the default), a
 * solve whose R has max|r_ii|/min|r_ii|·u > 1e-2 (u the unit roundoff of
 * `prec`), or whose refinement ends above tol, is refactored in FP64 and
 * refined again; cqr_lstsq_last_prec then reports CQR_PREC_F64.
 * iters / last_update (nullable) report the steps taken (over both
 * factorizations) and the last relative update.                          */
int cqr_lstsq(cqr_ctx* ctx, const double* A, int64_t lda, int64_t m, int64_t n, const double* B,
              int64_t ldb, int64_t nrhs, double* X, int64_t ldx, int64_t b, int prec, int max_iter,
              double tol, int* iters, double* last_update);
/* The solve and FP64 refinement of cqr_lstsq from factors already made:
 * YR holds [Y\R] of A (m×n, ld ldyr) as cqr_batch_qr or QR_2D leave it,
 * Tt the Tᵀ blocks of its panels of width b (⌈n/b⌉ blocks of b·b; one
 * n×n block for a batched problem, b = n).  A is the FP64 operator of the
 * residuals.  Statuses as cqr_lstsq.  No reference counterpart: it lets a
 * server factor many requests in one batch and then solve each.          */
int cqr_lstsq_factored(cqr_ctx* ctx, const double* A, int64_t lda, int64_t m, int64_t n, const double* YR,
                       int64_t ldyr, const double* Tt, int64_t b, const double* B, int64_t ldb, int64_t nrhs,
                       double* X, int64_t ldx, int max_iter, double tol, int* iters, double* last_update);
size_t cqr_lstsq_factored_workspace_bytes(int64_t m, int64_t n, int64_t nrhs, int64_t b);
/* Escalation of stalled mixed-precision solves to FP64 (on by default).   */
int cqr_set_lstsq_escalate(cqr_ctx* ctx, int on);
/* Precision of the factorization that produced the last cqr_lstsq's X
 * (CQR_PREC_*), or −1 when that call failed before factoring.            */
int cqr_lstsq_last_prec(const cqr_ctx* ctx);

/* ---- batched tall-skinny QR (no reference counterpart: a batch of
 *      independent hh_recon_qr panels, alg/QR/hh_recon/hh_recon.cxx:129-321,
 *      each problem being one panel of width n_p) ------------------------
 * Problems p = 0..nprob−1 of m_p × n_p (m_p ≥ n_p, n_p ≤ 128) are packed
 * one after another, column-major with ld = m_p, at the offsets
 * cqr_batch_offsets returns; their Tᵀ blocks (n_p × n_p) likewise.      */
typedef struct cqr_batch cqr_batch;
int cqr_batch_create(cqr_ctx* ctx, cqr_batch** batch, int64_t nprob, const int64_t* m,
                     const int64_t* n);
int cqr_batch_destroy(cqr_ctx* ctx, cqr_batch* batch);
int cqr_batch_offsets(const cqr_batch* batch, int64_t* aoff, int64_t* toff, int64_t* total_a,
                      int64_t* total_t);
int cqr_batch_groups(const cqr_batch* batch);   /* problems grouped by n   */
int cqr_batch_group_info(const cqr_batch* batch, int group, int64_t* b, int64_t* nprob,
                         int64_t* leaf_rows, int64_t* leaves, int64_t* nodes);
int cqr_batch_set_leaf_rows(cqr_ctx* ctx, cqr_batch* batch, int group, int64_t leaf_rows);
/* A := [Y\R] of every problem, Tt := their Tᵀ.  dtype 0: FP64 buffers;
 * 1: FP32 buffers (computed in FP64 through a staging copy).  use_graph:
 * the first call with given buffers runs eagerly and captures the launch
 * sequence in a HIP graph; later calls with the same buffers replay it.  */
int cqr_batch_qr(cqr_ctx* ctx, cqr_batch* batch, void* A, void* Tt, int dtype, int use_graph);
/* The same results through one hh_recon_qr launch sequence per problem
 * (the per-problem loop the batched form is measured against).           */
int cqr_batch_qr_loop(cqr_ctx* ctx, cqr_batch* batch, double* A, double* Tt);
/* Autotuner: per group, times every leaf height on a copy of A and keeps
 * the fastest (reps timed runs after a warm-up).                         */
int cqr_batch_tune(cqr_ctx* ctx, cqr_batch* batch, const double* A, int reps);

/* ---- least-squares request server (no reference counterpart) ----------
 * nslots host threads each drive their own context and stream; a request
 * is admitted (FIFO) when a slot is idle and its workspace footprint fits
 * the HBM budget.  A, B and X live in device memory and must stay valid
 * until the request completes.                                           */
enum cqr_tier { CQR_TIER_F64 = 0, CQR_TIER_MIXED = 1 };
typedef struct cqr_request {
  const double* A;
  int64_t lda, m, n;
  const double* B;
  int64_t ldb, nrhs;
  double* X;
  int64_t ldx;
  int tier;          /* CQR_TIER_F64: FP64 to 1e-12; CQR_TIER_MIXED: FP64
                        refinement to 1e-10, with the FP32 trailing update
                        when n ≥ 1024 (narrower: FP64 factors)            */
} cqr_request;
typedef struct cqr_result {
  int64_t id;
  int status, iters, slot;
  int prec;          /* factorization precision of X (CQR_PREC_*): a mixed
                        request escalated to FP64 reports CQR_PREC_F64     */
  int batch;         /* requests factored together with it (1: alone)      */
  double last_update;
  int64_t t_submit_ns, t_start_ns, t_end_ns;   /* steady clock            */
  int64_t bytes;                               /* admitted footprint      */
} cqr_result;
/* Least squares for many requests at once (n ≤ 128, nrhs ≤ 16 each): all
 * of them factored in FP64 by one batched TSQR + Householder reconstruction
 * (the launches of cqr_batch_qr, per width group), then each solved and
 * refined in FP64 as cqr_lstsq does, to tol within max_iter steps.  A, B
 * are read and X written as in cqr_lstsq; the tier field is not used.
 * status[p] (and iters[p], last_update[p] when non-null) report request p
 * with cqr_lstsq's statuses.  The context's arena holds everything
 * (cqr_lstsq_batched_workspace_bytes; 0 for requests it cannot take).
 * Each request's X depends only on that request.  No reference
 * counterpart: it is how the server below batches solves.                */
int cqr_lstsq_batched(cqr_ctx* ctx, int64_t nprob, const cqr_request* reqs, int max_iter, double tol,
                      int* status, int* iters, double* last_update);
size_t cqr_lstsq_batched_workspace_bytes(int64_t nprob, const cqr_request* reqs);
typedef struct cqr_server cqr_server;
/* Workspace bytes one least-squares solve needs (its admission footprint). */
size_t cqr_lstsq_workspace_bytes(int64_t m, int64_t n, int64_t nrhs, int64_t b);
/* Each slot drives its own stream.  HIP maps a process's streams onto a
 * few hardware queues (GPU_MAX_HW_QUEUES, four by default, read when HIP
 * starts): give the process at least nslots + 4 so that the slots'
 * streams do not share queues and serialise.                             */
int cqr_server_create(cqr_server** server, int device, int nslots, size_t hbm_budget);
/* Waits for the requests already running; queued requests that have not
 * started finish as CQR_ERR_CANCELLED (a thread waiting for one returns). */
int cqr_server_destroy(cqr_server* server);
/* One-panel requests (n ≤ 128, nrhs ≤ 16) of one tier and one width are
 * factored in batches of up to max_batch (default 32; 1 turns batching
 * off) through cqr_lstsq_batched.  Applies to requests submitted
 * afterwards.                                                           */
int cqr_server_set_batching(cqr_server* server, int max_batch);
/* Returns the request id (≥ 0) or −status (e.g. −CQR_ERR_NOMEM when the
 * footprint exceeds the whole budget).                                   */
int64_t cqr_server_submit(cqr_server* server, const cqr_request* request);
/* Blocks until request id finishes and consumes its result; returns
 * CQR_ERR_INVALID at once for an id never submitted or already collected
 * (by a wait or a drain).                                                */
int cqr_server_wait(cqr_server* server, int64_t id, cqr_result* result);
/* Moves up to max finished results (in completion order) into out.      */
int cqr_server_drain(cqr_server* server, cqr_result* out, int max, int* count);
int cqr_server_pending(cqr_server* server, int64_t* queued, size_t* bytes_in_use);

#ifdef __cplusplus
}
#endif
#endif /* CANDMC_QR_H */
