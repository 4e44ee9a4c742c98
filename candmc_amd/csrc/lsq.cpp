// Least-squares path on top of the 2D QR: Qᵀ/Q application with the stored
// compact-WY factors, block triangular solves with R, and the least-squares
// driver with iterative refinement on FP64 residuals.
//
// The Qᵀ application follows apply_tsqr_QT (alg/QR/tsqr/apply_tsqr_QT.cxx:
// 31-165) in spirit: the reflectors are applied panel by panel as blocked
// (WY) updates, here with the panel factors of QR_2D instead of the tree
// factors.  The solve and the refinement have no counterpart in the
// reference.  Refinement is the augmented-system scheme of Björck (1967):
// with r = b − A·x the residual, each step solves
//     [I  A] [dr]   [f]     f = b − r − A·x,
//     [Aᵀ 0] [dx] = [g],    g = −Aᵀ·r,
// through the computed factors (h = R⁻ᵀg, d = Qᵀf, dx = R⁻¹(d₁ − h),
// dr = Q·[h; d₂]) while f and g are formed in FP64.  Its fixed point is
// the exact least-squares solution, so a factorization computed with an
// FP32 trailing update converges to FP64 accuracy when κ(A)·u₃₂ < 1.
// Cheaper corrected semi-normal-equation steps (dx = R⁻¹R⁻ᵀAᵀr, no Q;
// Björck 1987) come first and are kept while they contract fast, which
// they do when κ(A)²·u of the factorization is small.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cqr_internal.h"

namespace cqr {

// B (m×nr) := Qᵀ·B (trans = 1) or Q·B (trans = 0), Q = H_1···H_k from the
// [Y\R] panels of a QR_2D of width b (ldy) and their Tᵀ blocks (b×b each,
// ld = the panel's width).
int apply_q_ws(cqr_ctx* c, const double* YR, int64_t ldy, const double* Tt, int64_t m, int64_t k, int64_t b,
               double* B, int64_t ldb, int64_t nr, int trans, double* ws) {
  const int64_t np = (k + b - 1) / b;
  if (nr <= 16) {
    // Few columns (solves, refinement): each panel is two HBM-bound thin
    // products over Y (its unit-lower top block masked inside them) and
    // one b×b×nr product with T
    double* scr = ws;
    double* Z = ws + (thin_gemm_scratch_doubles(m, b, (int)nr) + 31) / 32 * 32;
    double* Zp = Z + (b * nr + 31) / 32 * 32;
    for (int64_t q = 0; q < np; ++q) {
      const int64_t p = trans ? q : np - 1 - q;
      const int64_t j = p * b, bj = std::min(b, k - j), mj = m - j;
      if (mj <= 0) continue;
      const double* Yp = YR + j + j * ldy;
      double* Bp = B + j;
      const int bt = (int)std::min(bj, mj);
      if (nr <= 4 && bj <= 128 && mj >= 16384) {
        // Z = Yᵀ·B as slab partials, then their sum and Z' = Tᵀ·Z (Qᵀ) or
        // T·Z (Q) in one workgroup (it takes a whole CU's LDS: only for tall
        // panels, whose many slab partials pay for it); Tt holds Tᵀ (lower,
        // ld bj)
        CQR_TRY(thin_gemm(c, 1, mj, bj, 1.0, Yp, ldy, Bp, ldb, (int)nr, 0.0, nullptr, bj, scr, bt));
        const int64_t slab = thin_t_slab_rows(mj, bj);
        CQR_TRY(launch_slab_reduce_t(c, scr, (int)((mj + slab - 1) / slab), (int)bj, (int)nr, Tt + p * b * b, bj,
                                     trans ? 0 : 1, Zp, bj));
      } else {
        // Z = Yᵀ·B over all the panel's rows (its top block unit lower)
        CQR_TRY(thin_gemm(c, 1, mj, bj, 1.0, Yp, ldy, Bp, ldb, (int)nr, 0.0, Z, bj, scr, bt));
        // Z' = Tᵀ·Z (Qᵀ) or T·Z (Q); Tt holds Tᵀ (lower, ld bj)
        CQR_TRY(gemm_f64_ws(c, trans ? 0 : 1, 0, bj, nr, bj, 1.0, Tt + p * b * b, bj, SM_LOWER, Z, bj, SM_NONE, 0.0,
                            Zp, bj, nullptr));
      }
      // B −= Y·Z'
      CQR_TRY(thin_gemm(c, 0, mj, bj, -1.0, Yp, ldy, Zp, bj, (int)nr, 1.0, Bp, ldb, nullptr, bt));
    }
    return CQR_OK;
  }
  for (int64_t q = 0; q < np; ++q) {
    const int64_t p = trans ? q : np - 1 - q;
    const int64_t j = p * b, bj = std::min(b, k - j);
    if (m - j <= 0) continue;
    CQR_TRY(wy_update_ws(c, YR + j + j * ldy, ldy, 1, Tt + p * b * b, trans ? 0 : 1, B + j, ldb, m - j, nr,
                         (int)bj, ws, CQR_PREC_F64));
  }
  return CQR_OK;
}

size_t apply_q_ws_doubles(int64_t m, int64_t b, int64_t nr) {
  const size_t thin = nr <= 16 ? (thin_gemm_scratch_doubles(m, b, (int)nr) + 31) / 32 * 32 + 2 * ((b * nr + 31) / 32 * 32)
                               : 0;
  return std::max(wy_update_ws_doubles(m, nr, (int)b), thin) + 64;
}

// Inverses of the diagonal blocks of R (upper triangle of YR), 128 wide.
int r_block_inverses(cqr_ctx* c, const double* YR, int64_t ldy, int64_t n, double* Rinv) {
  return launch_tri_inverse_diag(c, YR, ldy, Rinv, n, 128);
}

// B(0:n, :) := R⁻¹·B (trans = 0) or R⁻ᵀ·B (trans = 1), blocked with the
// diagonal-block inverses; tmp holds 128×nr doubles.
int trsm_r_ws(cqr_ctx* c, const double* YR, int64_t ldy, int64_t n, const double* Rinv, double* B, int64_t ldb,
              int64_t nr, int trans, double* tmp) {
  const int64_t nb = (n + 127) / 128;
  // the one-launch solve from four blocks on: its workgroups each need a
  // whole CU, which beside other work (a server's other slots) costs more
  // than the two or three launches of a short solve
  if (nr <= 16 && nb >= 4 && nb <= CQR_TRSM_MAX_BLOCKS) return launch_trsm_chain(c, YR, ldy, n, Rinv, B, ldb, (int)nr, trans);
  if (nr <= 16) {
    // first diagonal block alone, then one launch per block step: the
    // remaining blocks take the new block into account and the next one
    // is finished by its own workgroup
    const int64_t p0 = trans ? 0 : nb - 1, j0 = p0 * 128, b0 = std::min<int64_t>(128, n - j0);
    CQR_TRY(launch_small_thin(c, trans ? 1 : 0, 0, b0, nr, b0, 1.0, Rinv + p0 * 128 * 128, b0, SM_UPPER, B + j0, ldb,
                              SM_NONE, 0.0, B + j0, ldb, PK_GEMM_SMALL));
    for (int64_t s = 1; s < nb; ++s) {
      if (trans) {   // block p = s−1 known; blocks s.. take it in, block s is finished
        CQR_TRY(launch_trsm_step(c, YR, ldy, n, Rinv, B, ldb, (int)nr, 1, (int)(s - 1), (int)s, (int)(nb - s),
                                 (int)s));
      } else {       // block p = nb−s known; blocks 0..p−1 take it in, block p−1 is finished
        const int p = (int)(nb - s);
        CQR_TRY(launch_trsm_step(c, YR, ldy, n, Rinv, B, ldb, (int)nr, 0, p, 0, p, p - 1));
      }
    }
    return CQR_OK;
  }
  for (int64_t q = 0; q < nb; ++q) {
    const int64_t p = trans ? q : nb - 1 - q;
    const int64_t j = p * 128, bj = std::min<int64_t>(128, n - j);
    const double* Ri = Rinv + p * 128 * 128;
    // x_p = R_pp⁻¹·y_p  (or R_pp⁻ᵀ·y_p); with few columns in place
    if (nr <= 16) {
      CQR_TRY(launch_small_thin(c, trans ? 1 : 0, 0, bj, nr, bj, 1.0, Ri, bj, SM_UPPER, B + j, ldb, SM_NONE, 0.0,
                                B + j, ldb, PK_GEMM_SMALL));
    } else {
      CQR_TRY(gemm_f64_ws(c, trans ? 1 : 0, 0, bj, nr, bj, 1.0, Ri, bj, SM_UPPER, B + j, ldb, SM_NONE, 0.0, tmp,
                          bj, nullptr));
      CQR_TRY(launch_copy2d(c, tmp, bj, B + j, ldb, bj, nr));
    }
    if (!trans && j > 0) {      // y(0:j) −= R(0:j, p)·x_p
      if (nr <= 16)
        CQR_TRY(thin_gemm(c, 0, j, bj, -1.0, YR + j * ldy, ldy, B + j, ldb, (int)nr, 1.0, B, ldb, nullptr));
      else
        CQR_TRY(gemm_f64_ws(c, 0, 0, j, nr, bj, -1.0, YR + j * ldy, ldy, SM_NONE, B + j, ldb, SM_NONE, 1.0, B,
                            ldb, nullptr));
    }
    if (trans && j + bj < n)    // y(j+bj:n) −= R(p, j+bj:n)ᵀ·x_p
      CQR_TRY(gemm_f64_ws(c, 1, 0, n - j - bj, nr, bj, -1.0, YR + j + (j + bj) * ldy, ldy, SM_NONE, B + j, ldb,
                          SM_NONE, 1.0, B + j + bj, ldb, nullptr));
  }
  return CQR_OK;
}

// R's block inverses and the vectors of the refinement.
static void refine_layout(int64_t m, int64_t n, int64_t nr, int64_t b, Sizer& s, LsqWS* w) {
  const int64_t nb = (n + 127) / 128;
  w->Ri = s.add(8 * nb * 128 * 128);
  w->R = s.add(8 * (size_t)m * nr);
  w->F = s.add(8 * (size_t)m * nr);
  w->G = s.add(8 * (size_t)n * nr);
  w->H = s.add(8 * (size_t)n * nr);
  w->Tmp = s.add(8 * 128 * nr);
  w->Q = s.add(8 * std::max({apply_q_ws_doubles(m, b, nr), thin_gemm_scratch_doubles(m, n, (int)nr),
                              thin_n_scratch_doubles(m, n, (int)nr)}));
  w->N = s.add(8 * 2 * nr);
}

// Workspace of the least-squares driver: the QR_2D arena followed by the
// factor copy, the panel factors and the refinement's buffers.
size_t lstsq_layout(int64_t m, int64_t n, int64_t nr, int64_t b, int64_t leaf_rows, Qr2dWS* o, LsqWS* w) {
  const int64_t np = (n + b - 1) / b;
  Sizer s;
  s.bytes = qr2d_ws(m, n, b, leaf_rows, o);
  w->Af = s.add(8 * (size_t)m * n);
  w->Tt = s.add(8 * np * b * b);
  refine_layout(m, n, nr, b, s, w);
  return s.bytes;
}

// The refinement's buffers alone (factors given by the caller).
size_t refine_ws(int64_t m, int64_t n, int64_t nr, int64_t b, LsqWS* w) {
  Sizer s;
  w->Af = w->Tt = 0;
  refine_layout(m, n, nr, b, s, w);
  return s.bytes;
}

}  // namespace cqr

using namespace cqr;

extern "C" {

size_t cqr_lstsq_workspace_bytes(int64_t m, int64_t n, int64_t nrhs, int64_t b) {
  if (m < 1 || n < 1 || m < n || nrhs < 1 || b < 1 || b > CQR_MAX_B) return 0;
  Qr2dWS o;
  LsqWS w;
  return lstsq_layout(m, n, nrhs, b, 0, &o, &w);
}

int cqr_apply_qt(cqr_ctx* c, const double* YR, int64_t ldy, const double* Tt, int64_t m, int64_t k, int64_t b,
                 double* B, int64_t ldb, int64_t nrhs, int trans) {
  CQR_CHECK_CTX(c);
  if (!YR || !Tt || !B || m < 1 || k < 1 || k > m || b < 1 || ldy < m || ldb < m || nrhs < 0)
    return set_err(c, CQR_ERR_INVALID, "apply_qt: bad arguments");
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "apply_qt: b > %d", CQR_MAX_B);
  if (nrhs == 0) return CQR_OK;
  CQR_TRY(ensure_ws(c, 8 * apply_q_ws_doubles(m, b, nrhs)));
  return apply_q_ws(c, YR, ldy, Tt, m, k, b, B, ldb, nrhs, trans ? 1 : 0, static_cast<double*>(c->ws));
}

int cqr_trsm_r(cqr_ctx* c, const double* YR, int64_t ldy, int64_t n, double* B, int64_t ldb, int64_t nrhs,
               int trans) {
  CQR_CHECK_CTX(c);
  if (!YR || !B || n < 1 || ldy < n || ldb < n || nrhs < 0) return set_err(c, CQR_ERR_INVALID, "trsm_r: bad arguments");
  if (nrhs == 0) return CQR_OK;
  const int64_t nb = (n + 127) / 128;
  Sizer s;
  const size_t oR = s.add(8 * nb * 128 * 128), oT = s.add(8 * 128 * nrhs);
  CQR_TRY(ensure_ws(c, s.bytes));
  char* base = static_cast<char*>(c->ws);
  double* Rinv = reinterpret_cast<double*>(base + oR);
  CQR_TRY(r_block_inverses(c, YR, ldy, n, Rinv));
  return trsm_r_ws(c, YR, ldy, n, Rinv, B, ldb, nrhs, trans ? 1 : 0, reinterpret_cast<double*>(base + oT));
}

}  // extern "C"

namespace cqr {

namespace {

// Unit roundoff of the trailing update's arithmetic.
double prec_unit(int prec) {
  return prec == CQR_PREC_F32 ? 0x1p-24 : prec == CQR_PREC_BF16 ? 0x1p-8 : 0x1p-53;
}

// Beyond this κ̂·u the mixed factorization is not worth refining: the
// stationary refinement contracts by about κ·u per step (and not at all
// from κ·u ≈ 1), so the FP64 factorization is taken at once.
constexpr double kEscalateKappaU = 1e-2;
enum { ATTEMPT_ESCALATE = -1 };

}  // namespace

// One factorization in `prec` and its refinement (lstsq_refine).
static int lstsq_attempt(cqr_ctx* c, const double* A, int64_t lda, int64_t m, int64_t n, const double* B, int64_t ldb,
                  int64_t nr, double* X, int64_t ldx, int64_t b, int prec, int max_iter, double tol, bool escalate,
                  char* base, const Qr2dWS& o, const LsqWS& w, LsqRun& out) {
  double* Af = reinterpret_cast<double*>(base + w.Af);
  double* Tt = reinterpret_cast<double*>(base + w.Tt);
  // factor a copy of A (the caller's A stays the FP64 reference operator)
  CQR_TRY(launch_copy2d(c, A, lda, Af, m, m, n));
  CQR_TRY(qr2d_run(c, Af, m, m, n, b, nullptr, 0, base, o, Tt, prec));
  return lstsq_refine(c, A, lda, m, n, B, ldb, nr, X, ldx, b, Af, m, Tt, prec, max_iter, tol, escalate, base, w,
                      out);
}

int lstsq_refine(cqr_ctx* c, const double* A, int64_t lda, int64_t m, int64_t n, const double* B, int64_t ldb,
                 int64_t nr, double* X, int64_t ldx, int64_t b, const double* Af, int64_t ldf, const double* Tt,
                 int prec, int max_iter, double tol, bool escalate, char* base, const LsqWS& w, LsqRun& out) {
  auto at = [&](size_t off) { return reinterpret_cast<double*>(base + off); };
  double *Rinv = at(w.Ri), *r = at(w.R), *f = at(w.F), *g = at(w.G), *h = at(w.H), *tmp = at(w.Tmp),
         *wsq = at(w.Q), *nrm = at(w.N);
  double* hn = c->h_pin;   // pinned: 2·nr ≤ 32 doubles
  // κ̂ = max|r_ii| / min|r_ii|, a lower bound on κ(A): copied out now and
  // read once x0 is formed in the workspace, before X is first written
  CQR_TRY(launch_diag_range(c, Af, ldf, n, nrm));
  CQR_HIP(c, hipMemcpyAsync(hn, nrm, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CQR_HIP(c, hipEventRecord(c->ev_host, c->stream));
  CQR_TRY(r_block_inverses(c, Af, ldf, n, Rinv));
  // x0 = R⁻¹·(Qᵀb)(0:n)
  CQR_TRY(launch_copy2d(c, B, ldb, f, m, m, nr));
  CQR_TRY(apply_q_ws(c, Af, ldf, Tt, m, n, b, f, m, nr, 1, wsq));
  CQR_TRY(trsm_r_ws(c, Af, ldf, n, Rinv, f, m, nr, 0, tmp));
  CQR_HIP(c, hipEventSynchronize(c->ev_host));
  if (!std::isfinite(hn[0]) || !std::isfinite(hn[1])) return set_err(c, CQR_ERR_NONFINITE, "lstsq: R is not finite");
  if (escalate && prec != CQR_PREC_F64 && !(hn[1] > 0.0 && hn[0] / hn[1] * prec_unit(prec) <= kEscalateKappaU))
    return ATTEMPT_ESCALATE;
  CQR_TRY(launch_copy2d(c, f, m, X, ldx, n, nr));
  if (max_iter == 0) return CQR_OK;
  // r = b − A·x
  CQR_TRY(launch_copy2d(c, B, ldb, r, m, m, nr));
  CQR_TRY(thin_gemm(c, 0, m, n, -1.0, A, lda, X, ldx, (int)nr, 1.0, r, m, wsq));
  // max over the columns of ‖d_j‖ / ‖x_j‖ (one host synchronisation); NaN
  // if any column's update or solution norm is NaN
  auto rel_update = [&](const double* d, int64_t ldd, double* res) -> int {
    CQR_TRY(launch_colnorm(c, d, ldd, n, (int)nr, nrm));
    CQR_TRY(launch_colnorm(c, X, ldx, n, (int)nr, nrm + nr));
    CQR_HIP(c, hipMemcpyAsync(hn, nrm, sizeof(double) * 2 * nr, hipMemcpyDeviceToHost, c->stream));
    CQR_HIP(c, hipStreamSynchronize(c->stream));
    double rel = 0.0;
    for (int64_t j = 0; j < nr; ++j) {
      if (std::isnan(hn[j]) || std::isnan(hn[nr + j])) {
        rel = NAN;
        break;
      }
      rel = std::max(rel, hn[j] / std::max(hn[nr + j], 1e-300));
    }
    *res = rel;
    return CQR_OK;
  };
  double prev_rel = 0.0;
  bool sne = true;   // semi-normal-equation corrections while they contract fast
  int aug = 0;       // augmented-system steps taken
  for (int it = 1; it <= max_iter; ++it) {
    if (sne) {
      // dx = R⁻¹·R⁻ᵀ·(Aᵀ·r): the corrected semi-normal equations with the
      // computed R (no Q application).  Their fixed point is also the exact
      // least-squares solution; they contract by about κ²·u of the
      // factorization, so they are used while each update shrinks by at
      // least 8× (and the first is below half the solution).
      CQR_TRY(thin_gemm(c, 1, m, n, 1.0, A, lda, r, m, (int)nr, 0.0, g, n, wsq));
      CQR_TRY(trsm_r_ws(c, Af, ldf, n, Rinv, g, n, nr, 1, tmp));
      CQR_TRY(trsm_r_ws(c, Af, ldf, n, Rinv, g, n, nr, 0, tmp));
      double rel;
      CQR_TRY(rel_update(g, n, &rel));
      if (rel <= 0.5 && (it == 1 || rel <= prev_rel / 8)) {
        CQR_TRY(launch_axpby(c, n, nr, 1.0, g, n, 1.0, X, ldx));
        CQR_TRY(launch_copy2d(c, B, ldb, r, m, m, nr));
        CQR_TRY(thin_gemm(c, 0, m, n, -1.0, A, lda, X, ldx, (int)nr, 1.0, r, m, wsq));
        out.iters += 1;
        out.rel = rel;
        if (rel <= tol) return CQR_OK;
        prev_rel = rel;
        continue;
      }
      if (std::isnan(rel)) {
        out.rel = rel;
        return set_err(c, CQR_ERR_NONFINITE, "lstsq: the refinement produced NaN");
      }
      sne = false;   // not contracting fast enough: the augmented system from here
    }
    ++aug;
    // f = b − r − A·x,  g = −Aᵀ·r   (FP64)
    CQR_TRY(launch_copy2d(c, B, ldb, f, m, m, nr));
    CQR_TRY(launch_axpby(c, m, nr, -1.0, r, m, 1.0, f, m));
    CQR_TRY(thin_gemm(c, 0, m, n, -1.0, A, lda, X, ldx, (int)nr, 1.0, f, m, wsq));
    CQR_TRY(thin_gemm(c, 1, m, n, -1.0, A, lda, r, m, (int)nr, 0.0, g, n, wsq));
    // h = R⁻ᵀ·g;  d = Qᵀ·f;  dx = R⁻¹·(d₁ − h);  dr = Q·[h; d₂]
    CQR_TRY(trsm_r_ws(c, Af, ldf, n, Rinv, g, n, nr, 1, tmp));               // g := h
    CQR_TRY(apply_q_ws(c, Af, ldf, Tt, m, n, b, f, m, nr, 1, wsq));           // f := d
    CQR_TRY(launch_copy2d(c, f, m, h, n, n, nr));                            // h := d₁
    CQR_TRY(launch_axpby(c, n, nr, -1.0, g, n, 1.0, h, n));                  // h := d₁ − h
    CQR_TRY(trsm_r_ws(c, Af, ldf, n, Rinv, h, n, nr, 0, tmp));                // h := dx
    CQR_TRY(launch_copy2d(c, g, n, f, m, n, nr));                            // f := [h; d₂]
    CQR_TRY(apply_q_ws(c, Af, ldf, Tt, m, n, b, f, m, nr, 0, wsq));           // f := dr
    CQR_TRY(launch_axpby(c, n, nr, 1.0, h, n, 1.0, X, ldx));
    CQR_TRY(launch_axpby(c, m, nr, 1.0, f, m, 1.0, r, m));
    // stop when every column's update is below tol relative to the solution
    double rel;
    CQR_TRY(rel_update(h, n, &rel));
    out.iters += 1;
    out.rel = rel;
    if (std::isnan(rel)) return set_err(c, CQR_ERR_NONFINITE, "lstsq: the refinement produced NaN");
    if (rel <= tol) return CQR_OK;
    // the updates stopped shrinking: the attainable accuracy (about κ·ε of
    // the FP64 residual) is reached above tol, or κ·u of the factorization
    // is ≥ 1
    if (aug > 1 && rel > 0.9 * prev_rel) break;
    prev_rel = rel;
  }
  return set_err(c, CQR_ERR_NOCONVERGE, "lstsq: the last relative update %.3g is above tol %.3g after %d steps",
                 out.rel, tol, out.iters);
}


}  // namespace cqr

extern "C" {

size_t cqr_lstsq_factored_workspace_bytes(int64_t m, int64_t n, int64_t nrhs, int64_t b) {
  if (m < 1 || n < 1 || m < n || nrhs < 1 || b < 1 || b > CQR_MAX_B) return 0;
  LsqWS w;
  return refine_ws(m, n, nrhs, b, &w);
}

int cqr_lstsq_factored(cqr_ctx* c, const double* A, int64_t lda, int64_t m, int64_t n, const double* YR,
                       int64_t ldyr, const double* Tt, int64_t b, const double* B, int64_t ldb, int64_t nrhs,
                       double* X, int64_t ldx, int max_iter, double tol, int* iters, double* last_update) {
  CQR_CHECK_CTX(c);
  if (iters) *iters = 0;
  if (last_update) *last_update = 0.0;
  if (!A || !YR || !Tt || !B || !X || m < 1 || n < 1 || m < n || lda < m || ldyr < m || ldb < m || ldx < n ||
      nrhs < 1 || b < 1 || max_iter < 0 || !(tol >= 0.0))
    return set_err(c, CQR_ERR_INVALID, "lstsq_factored: bad arguments");
  if (nrhs > 16) return set_err(c, CQR_ERR_UNSUPPORTED, "lstsq_factored: at most 16 right-hand sides");
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "lstsq_factored: b > %d", CQR_MAX_B);
  LsqWS w;
  CQR_TRY(ensure_ws(c, refine_ws(m, n, nrhs, b, &w)));
  LsqRun run;
  const int st = lstsq_refine(c, A, lda, m, n, B, ldb, nrhs, X, ldx, b, YR, ldyr, Tt, CQR_PREC_F64, max_iter, tol,
                              false, static_cast<char*>(c->ws), w, run);
  if (iters) *iters = run.iters;
  if (last_update) *last_update = run.rel;
  return st;
}

int cqr_set_lstsq_escalate(cqr_ctx* c, int on) {
  CQR_CHECK_CTX(c);
  c->lsq_escalate = on ? 1 : 0;
  return CQR_OK;
}

int cqr_lstsq_last_prec(const cqr_ctx* c) { return c ? c->lsq_last_prec : -1; }

int cqr_lstsq(cqr_ctx* c, const double* A, int64_t lda, int64_t m, int64_t n, const double* B, int64_t ldb,
              int64_t nrhs, double* X, int64_t ldx, int64_t b, int prec, int max_iter, double tol, int* iters,
              double* last_update) {
  CQR_CHECK_CTX(c);
  if (iters) *iters = 0;
  if (last_update) *last_update = 0.0;
  c->lsq_last_prec = -1;
  if (!A || !B || !X || m < 1 || n < 1 || m < n || lda < m || ldb < m || ldx < n || nrhs < 1 || b < 1 ||
      max_iter < 0 || !(tol >= 0.0))
    return set_err(c, CQR_ERR_INVALID, "lstsq: bad arguments (needs m >= n >= 1, nrhs >= 1, tol >= 0)");
  if (nrhs > 16) return set_err(c, CQR_ERR_UNSUPPORTED, "lstsq: at most 16 right-hand sides");
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "lstsq: b > %d", CQR_MAX_B);
  if (prec != CQR_PREC_F64 && prec != CQR_PREC_F32 && prec != CQR_PREC_BF16) return set_err(c, CQR_ERR_UNSUPPORTED, "lstsq: precision %d", prec);
  CQR_TRY(finite_check(c, A, lda, m, n));
  Qr2dWS o;
  LsqWS w;
  CQR_TRY(ensure_ws(c, lstsq_layout(m, n, nrhs, b, c->leaf_rows, &o, &w)));
  char* base = static_cast<char*>(c->ws);
  const bool esc = c->lsq_escalate && prec != CQR_PREC_F64;
  LsqRun run;
  int st = lstsq_attempt(c, A, lda, m, n, B, ldb, nrhs, X, ldx, b, prec, max_iter, tol, esc, base, o, w, run);
  int used = prec;
  // a mixed factorization ruled out by R's diagonal, or whose refinement
  // ended above tol: refactor in FP64 and refine again
  if (esc && (st == ATTEMPT_ESCALATE || st == CQR_ERR_NOCONVERGE)) {
    used = CQR_PREC_F64;
    st = lstsq_attempt(c, A, lda, m, n, B, ldb, nrhs, X, ldx, b, CQR_PREC_F64, max_iter, tol, false, base, o, w, run);
  }
  if (iters) *iters = run.iters;
  if (last_update) *last_update = run.rel;
  c->lsq_last_prec = used;
  return st;
}

}  // extern "C"
