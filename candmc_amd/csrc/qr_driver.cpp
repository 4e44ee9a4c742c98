// Host drivers of the CA-QR hot path: the C-ABI entry points declared in
// include/candmc_qr.h.  Each follows the reference function it replaces;
// MPI hand-offs become stream-ordered kernel launches over in-HBM buffers.
#include <cstring>

#include "cqr_internal.h"

namespace cqr {

int launch_tril_halve(cqr_ctx* c, const double* G, int b, double* T, int64_t ldt);
int launch_copy2d(cqr_ctx* c, const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m,
                  int64_t n);

// ---------------------------------------------------------------------------
// Panel workspace: every buffer TSQR + HR needs for one m×b panel.
// ---------------------------------------------------------------------------
struct PanelWS {
  size_t leafR, leafT, leafTau, nodeR, nodeV2, nodeT, nodeTau, C, Y, Uinv, Winv, LinvT, signs, Wtmp;
};

static PanelWS panel_layout(const TsqrPlan& p, Sizer& s) {
  const size_t b = p.b, bb = b * b, L = p.nleaves, nn = std::max<int64_t>(p.nnodes, 1);
  PanelWS w;
  w.leafR = s.add(8 * L * bb);
  w.leafT = s.add(8 * L * bb);
  w.leafTau = s.add(8 * L * b);
  w.nodeR = s.add(8 * nn * bb);
  w.nodeV2 = s.add(8 * nn * bb);
  w.nodeT = s.add(8 * nn * bb);
  w.nodeTau = s.add(8 * nn * b);
  w.C = s.add(8 * L * bb);
  w.Y = s.add(8 * (size_t)p.m * b);
  w.Uinv = s.add(8 * 2 * bb);   // [U⁻¹ | U⁻¹·Tᵀ], b×2b
  w.Winv = s.add(8 * bb);
  w.LinvT = s.add(8 * bb);
  w.signs = s.add(8 * b);
  w.Wtmp = s.add(8 * bb);
  return w;
}

static BgemmArgs bg(int m, int n, int k, double alpha, double beta) {
  BgemmArgs g{};
  g.m = m; g.n = n; g.k = k; g.alpha = alpha; g.beta = beta;
  return g;
}

size_t panel_ws_bytes(int64_t m, int64_t b, int64_t nleaves, int64_t leaf_rows) {
  TsqrPlan p;
  if (make_plan(m, b, nleaves, leaf_rows, p) != CQR_OK) return 0;
  Sizer s;
  panel_layout(p, s);
  return s.bytes;
}

// TSQR + Householder reconstruction of one panel (hh_recon.cxx:129-321).
// Outputs: A := [Y\R], W (b×b), and optionally T, T⁻ᵀ and Tᵀ (all ld b).
// With Tt, Ye and YT all given (m > b), also the explicit unit-lower Y in Ye
// and Y·Tᵀ in YT (both m×b, ld m), the lower parts from the same pass over
// Q1 that forms Y: [Y | Y·Tᵀ] = Q1·[U⁻¹ | U⁻¹·Tᵀ].
int hh_recon_panel(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t b, double* W, double* T,
                   double* invT_T, double* Tt, int64_t nleaves, char* base, int64_t* signs_out, double* Ye,
                   double* YT) {
  TsqrPlan p;
  int s = make_plan(m, b, nleaves, c->leaf_rows, p);
  if (s != CQR_OK) return set_err(c, s, "hh_recon_qr: cannot plan m=%lld b=%lld", (long long)m, (long long)b);
  Sizer sz;
  const PanelWS w = panel_layout(p, sz);
  auto at = [&](size_t off) { return reinterpret_cast<double*>(base + off); };
  const int bi = (int)b, L = (int)p.nleaves;
  const int64_t bb = b * b;
  double* leafR = at(w.leafR);
  double* leafT = at(w.leafT);
  double* nodeV2 = at(w.nodeV2);
  double* nodeT = at(w.nodeT);
  // K1 + K2: leaves (bitree_tsqr.cxx:156) and the reduction tree, all levels
  // pipelined (bitree_tsqr.cxx:172-232), fused into one launch where possible
  CQR_TRY(leaf_tree(c, A, lda, m, L, bi, leafR, leafT, at(w.leafTau), at(w.nodeR), nodeV2, nodeT, at(w.nodeTau),
                    (int)p.nnodes));
  const double* Rroot = p.nnodes > 0 ? at(w.nodeR) + (p.nnodes - 1) * bb : leafR;
  // K3: construct_Q1 (construct_tsqr_Q.cxx:34-242) into the Y scratch
  double* Y = at(w.Y);
  CQR_TRY(launch_q1(c, nodeV2, nodeT, leafT, A, lda, m, at(w.C), Y, m, L, (int)p.nnodes, bi));
  if (c->gate_at) CQR_HIP(c, hipEventRecord(c->ev_gate, c->stream));
  // K4: LU(Q1_top − S), R := S·R, W = U·S (hh_recon.cxx:282-308)
  double* Wd = W ? W : at(w.Wtmp);
  int64_t* sg = signs_out ? signs_out : reinterpret_cast<int64_t*>(base + w.signs);
  double* Uinv = at(w.Uinv);
  const bool expl = m > b && Tt && Ye && YT;
  // at b = 128 the HR core also forms Tᵀ (and, for the explicit Y, U⁻¹·Tᵀ
  // and the top b rows of Y and Y·Tᵀ) in the same launch
  const bool fused = Tt && b == 128;
  CQR_TRY(hr_core(c, Y, m, Rroot, A, lda, Wd, Uinv, at(w.Winv), at(w.LinvT), sg, bi, fused ? Tt : nullptr,
                  fused && expl ? Uinv + bb : nullptr, fused && expl ? Ye : nullptr, fused && expl ? YT : nullptr,
                  m));
  // T = −W·L⁻ᵀ, Tᵀ = −L⁻¹·Wᵀ, T⁻ᵀ = −W⁻ᵀ·L (W = −T·Y1ᵀ)
  if (Tt && !fused)
    CQR_TRY(gemm_f64_ws(c, 1, 1, b, b, b, -1.0, at(w.LinvT), b, SM_UNIT_UPPER, Wd, b, SM_UPPER, 0.0, Tt, b,
                        nullptr));
  // Y = Q1·U⁻¹ below the top b rows (hh_recon.cxx:294-301)
  // (U⁻¹ is stored with explicit zeros below its diagonal: a plain GEMM)
  if (expl) {
    if (!fused) {
      CQR_TRY(gemm_f64_ws(c, 0, 0, b, b, b, 1.0, Uinv, b, SM_NONE, Tt, b, SM_NONE, 0.0, Uinv + bb, b, nullptr));
      CQR_TRY(launch_copy_unit_lower(c, A, lda, Ye, m, b, b));
      // Y1·Tᵀ = −Wᵀ exactly (W = −T·Y1ᵀ); formed as a product like the rest
      CQR_TRY(gemm_f64_ws(c, 0, 0, b, b, b, 1.0, A, lda, SM_UNIT_LOWER, Tt, b, SM_NONE, 0.0, YT, m, nullptr));
    }
    CQR_TRY(update_gemm_split(c, m - b, 2 * b, b, b, Y + b, m, Uinv, b, A + b, lda, Ye + b, m, YT + b, m,
                              PK_GEMM_SMALL));
  } else if (m > b) {
    CQR_TRY(update_gemm(c, m - b, b, b, 1.0, Y + b, m, Uinv, b, 0.0, A + b, lda, PK_GEMM_SMALL, 0));
  }
  if (T)
    CQR_TRY(gemm_f64_ws(c, 0, 0, b, b, b, -1.0, Wd, b, SM_UPPER, at(w.LinvT), b, SM_UNIT_UPPER, 0.0, T, b,
                        nullptr));
  if (invT_T)
    CQR_TRY(gemm_f64_ws(c, 1, 0, b, b, b, -1.0, at(w.Winv), b, SM_UPPER, A, lda, SM_UNIT_LOWER, 0.0, invT_T, b,
                        nullptr));
  return CQR_OK;
}

int finite_check(cqr_ctx* c, const double* A, int64_t lda, int64_t m, int64_t n) {
  if (!c->check_finite) return CQR_OK;
  CQR_HIP(c, hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
  CQR_TRY(launch_finite_scan(c, A, lda, m, n));
  int h = 0;
  CQR_HIP(c, hipMemcpyAsync(&h, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  CQR_HIP(c, hipStreamSynchronize(c->stream));
  if (h) return set_err(c, CQR_ERR_NONFINITE, "input holds NaN or Inf");
  return CQR_OK;
}

// Tᵀ from T⁻ᵀ (lower b×b, ld ldi): lower-triangular inverse.
static int Tt_from_invTT(cqr_ctx* c, const double* invTT, int64_t ldi, double* Tt, int b) {
  return launch_tri_inverse(c, invTT, ldi, Tt, b, b, 0, 0);
}

// T⁻ᵀ = −W⁻ᵀ·Y1 (hh_recon.cxx:26-31); Y1 given with its mask.
static int invTT_from_W(cqr_ctx* c, const double* W, int b, const double* Y1, int64_t ldy, int smY,
                        double* Winv_scratch, double* out) {
  CQR_TRY(launch_tri_inverse(c, W, b, Winv_scratch, b, b, 1, 0));
  BgemmArgs g = bg(b, b, b, -1.0, 0.0);
  g.A = bmat(Winv_scratch, b); g.ta = 1; g.smA = SM_UPPER;
  g.B = bmat(Y1, ldy); g.smB = smY;
  g.C = bmat(out, b);
  return bgemm(c, g, 1, b);
}

}  // namespace cqr

using namespace cqr;

extern "C" {

int64_t cqr_tree_data_doubles(int64_t m, int64_t b, int64_t nleaves) {
  TsqrPlan p;
  if (make_plan(m, b, nleaves, 0, p) != CQR_OK) return -1;
  return p.nleaves * b * b + p.nnodes * (2 * b * b + b);
}

int cqr_hh_recon_qr_ex(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t b, double* W,
                       double* T, double* invT_T, int64_t nleaves) {
  CQR_CHECK_CTX(c);
  if (!A || !W || b < 1 || m < b || lda < m || nleaves < 0)
    return set_err(c, CQR_ERR_INVALID, "hh_recon_qr: bad arguments (m=%lld b=%lld lda=%lld)",
                   (long long)m, (long long)b, (long long)lda);
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "hh_recon_qr: b=%lld > %d", (long long)b, CQR_MAX_B);
  CQR_TRY(finite_check(c, A, lda, m, b));
  const size_t need = panel_ws_bytes(m, b, nleaves, c->leaf_rows);
  if (!need) return set_err(c, CQR_ERR_UNSUPPORTED, "hh_recon_qr: no leaf plan");
  CQR_TRY(ensure_ws(c, need));
  return hh_recon_panel(c, A, lda, m, b, W, T, invT_T, nullptr, nleaves, (char*)c->ws, nullptr);
}

int cqr_hh_recon_qr(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t b, double* W, int64_t nleaves) {
  return cqr_hh_recon_qr_ex(c, A, lda, m, b, W, nullptr, nullptr, nleaves);
}

}  // extern "C"

namespace cqr {

size_t local_tsqr_ws_bytes(int64_t m, int64_t b, int64_t leaf_rows) {
  const int64_t bb = b * b;
  if (m <= leaf_capacity_rows(b)) return 8 * 2 * bb + 256;
  Sizer s;
  s.add(panel_ws_bytes(m, b, 0, leaf_rows));
  s.add(8 * bb);
  s.add(8 * bb);
  return s.bytes;
}

// dgeqrf of one m×b block (bitree_tsqr.cxx:48-71) on the scratch at base
// (local_tsqr_ws_bytes): one leaf in registers, or, for taller blocks,
// TSQR + HR, which gives the same [V\R]; tau = diag(T).
int local_tsqr_ws(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t b, double* tau, double* T, char* base) {
  const int64_t bb = b * b;
  if (m <= leaf_capacity_rows(b)) {
    double* R = reinterpret_cast<double*>(base);
    double* Tw = T ? T : R + bb;
    return leaf_qr(c, A, lda, m, 1, (int)b, R, Tw, tau);
  }
  Sizer s;
  const size_t pw = s.add(panel_ws_bytes(m, b, 0, c->leaf_rows));
  const size_t tw = s.add(8 * bb), ww = s.add(8 * bb);
  double* Tw = T ? T : reinterpret_cast<double*>(base + tw);
  CQR_TRY(hh_recon_panel(c, A, lda, m, b, reinterpret_cast<double*>(base + ww), Tw, nullptr, nullptr, 0,
                         base + pw, nullptr));
  return launch_diag(c, Tw, b, tau, (int)b);
}

}  // namespace cqr

extern "C" {

int cqr_local_tsqr(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t b, double* tau, double* T) {
  CQR_CHECK_CTX(c);
  if (!A || !tau || b < 1 || m < b || lda < m) return set_err(c, CQR_ERR_INVALID, "local_tsqr: bad arguments");
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "local_tsqr: b > %d", CQR_MAX_B);
  CQR_TRY(finite_check(c, A, lda, m, b));
  CQR_TRY(ensure_ws(c, local_tsqr_ws_bytes(m, b, c->leaf_rows)));
  return local_tsqr_ws(c, A, lda, m, b, tau, T, (char*)c->ws);
}

int cqr_tree_tsqr(cqr_ctx* c, double* A2b, int64_t b, double* T) {
  CQR_CHECK_CTX(c);
  if (!A2b || b < 1) return set_err(c, CQR_ERR_INVALID, "tree_tsqr: bad arguments");
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "tree_tsqr: b > %d", CQR_MAX_B);
  const int64_t bb = b * b;
  CQR_TRY(ensure_ws(c, 8 * (5 * bb + b) + 1024));
  double* Rl = static_cast<double*>(c->ws);   // [R1 | R2], ld b
  double* V2 = Rl + 2 * bb;
  double* Tf = V2 + bb;
  double* tau = Tf + bb;
  double* Tb = tau + b;
  CQR_TRY(launch_copy2d(c, A2b, 2 * b, Rl, b, b, b));
  CQR_TRY(launch_copy2d(c, A2b + b, 2 * b, Rl + bb, b, b, b));
  CQR_TRY(tree_level(c, Rl, 2, 0, (int)b, V2, Tf, tau));
  CQR_TRY(launch_copy2d(c, Rl, b, A2b, 2 * b, b, b));
  CQR_TRY(launch_copy2d(c, V2, b, A2b + b, 2 * b, b, b));
  if (T) {
    const int nb = (int)std::min<int64_t>(b, 16);
    CQR_TRY(launch_blocked_T(c, Tf, Tb, (int)b, nb));
    CQR_TRY(launch_copy2d(c, Tb, nb, T, nb, nb, b));
  }
  return CQR_OK;
}

int cqr_bitree_tsqr(cqr_ctx* c, double* A, int64_t lda, double* R, double* tau, int64_t m, int64_t b,
                    int64_t nleaves, int output_Y, double* tree_data) {
  CQR_CHECK_CTX(c);
  if (!A || !R || !tau || b < 1 || m < b || lda < m || nleaves < 0)
    return set_err(c, CQR_ERR_INVALID, "bitree_tsqr: bad arguments");
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "bitree_tsqr: b > %d", CQR_MAX_B);
  // no tree_data: the reference's per-rank semantics over nleaves ranks
  if (!tree_data) return bitree_ranks(c, A, lda, R, tau, m, b, std::max<int64_t>(nleaves, 1), output_Y);
  TsqrPlan p;
  int s = make_plan(m, b, nleaves, 0, p);
  if (s != CQR_OK) return set_err(c, s, "bitree_tsqr: cannot plan");
  CQR_TRY(finite_check(c, A, lda, m, b));
  const int64_t bb = b * b, L = p.nleaves, nn = p.nnodes;
  CQR_TRY(ensure_ws(c, 8 * ((L + nn) * bb) + 512));
  double* leafR = static_cast<double*>(c->ws);
  double* nodeR = leafR + L * bb;
  double* leafT = tree_data;
  double* nodeV2 = leafT + L * bb;
  double* nodeT = nodeV2 + nn * bb;
  double* nodeTau = nodeT + nn * bb;
  CQR_TRY(leaf_tree(c, A, lda, m, (int)L, (int)b, leafR, leafT, tau, nodeR, nodeV2, nodeT, nodeTau, (int)nn));
  return launch_copy2d(c, nn > 0 ? nodeR + (nn - 1) * bb : leafR, b, R, b, b, b);
}

int cqr_construct_Q1(cqr_ctx* c, const double* Y, int64_t lda_Y, const double* tau, double* Q1,
                     int64_t lda_Q, int64_t m, int64_t b, int64_t k, int64_t nleaves,
                     const double* tree_data) {
  CQR_CHECK_CTX(c);
  // no tree_data: the one-rank dorgqr branch (construct_tsqr_Q.cxx:87-102)
  if (!tree_data) return cqr_construct_Q1_id(c, Y, lda_Y, tau, Q1, lda_Q, m, b, k, nleaves, nullptr, nullptr, b);
  if (!Y || !Q1 || b < 1 || m < b || lda_Y < m || lda_Q < m)
    return set_err(c, CQR_ERR_INVALID, "construct_Q1: bad arguments");
  if (k != b) return set_err(c, CQR_ERR_UNSUPPORTED, "construct_Q1: only k == b is supported");
  TsqrPlan p;
  int s = make_plan(m, b, nleaves, 0, p);
  if (s != CQR_OK) return set_err(c, s, "construct_Q1: cannot plan");
  const int64_t bb = b * b, L = p.nleaves, nn = p.nnodes;
  Sizer sz;
  const size_t oC = sz.add(8 * L * bb);
  CQR_TRY(ensure_ws(c, sz.bytes));
  char* base = (char*)c->ws;
  auto at = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
  const double* leafT = tree_data;
  const double* nodeV2 = leafT + L * bb;
  const double* nodeT = nodeV2 + nn * bb;
  return launch_q1(c, nodeV2, nodeT, leafT, Y, lda_Y, m, at(oC), Q1, lda_Q, (int)L, (int)nn, (int)b);
}

int cqr_signed_nlu(cqr_ctx* c, double* A, int64_t lda, double* R, int64_t lda_R, int64_t b, int64_t* signs) {
  CQR_CHECK_CTX(c);
  if (!A || b < 1 || lda < b || (R && lda_R < b)) return set_err(c, CQR_ERR_INVALID, "signed_NLU: bad arguments");
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "signed_NLU: b > %d", CQR_MAX_B);
  return launch_signed_nlu(c, A, lda, R, lda_R, signs, (int)b);
}

int cqr_recursive_nlu(cqr_ctx* c, double* A, int64_t lda, double* R, int64_t lda_R, int64_t b, int64_t* signs) {
  // The recursion of hh_recon.cxx:80-111 is a cache blocking of the same
  // factorization; one workgroup holds the whole block in LDS instead.
  return cqr_signed_nlu(c, A, lda, R, lda_R, b, signs);
}

int cqr_compute_invT_from_W(cqr_ctx* c, const double* W, int64_t b, double* invT) {
  CQR_CHECK_CTX(c);
  if (!W || !invT || b < 1) return set_err(c, CQR_ERR_INVALID, "compute_invT_from_W: bad arguments");
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "compute_invT_from_W: b > %d", CQR_MAX_B);
  const int64_t bb = b * b;
  CQR_TRY(ensure_ws(c, 8 * 2 * bb + 256));
  double* Winv = static_cast<double*>(c->ws);
  double* out = Winv + bb;
  CQR_TRY(invTT_from_W(c, W, (int)b, invT, b, SM_NONE, Winv, out));
  return launch_copy2d(c, out, b, invT, b, b, b);
}

int cqr_compute_invT_from_Y(cqr_ctx* c, const double* Y, int64_t lda_Y, double* T, int64_t lda_T, int64_t mb,
                            int64_t b) {
  CQR_CHECK_CTX(c);
  if (!Y || !T || b < 1 || mb < 0 || lda_Y < mb || lda_T < b)
    return set_err(c, CQR_ERR_INVALID, "compute_invT_from_Y: bad arguments");
  const int64_t bb = b * b;
  const size_t sd = gemm_f64_scratch_doubles(b, b, mb);
  Sizer s;
  const size_t og = s.add(8 * bb), op = s.add(8 * (sd + 1));
  CQR_TRY(ensure_ws(c, s.bytes));
  char* base = (char*)c->ws;
  double* G = reinterpret_cast<double*>(base + og);
  double* part = reinterpret_cast<double*>(base + op);
  if (mb > 0) {
    CQR_TRY(gemm_f64_ws(c, 1, 0, b, b, mb, 1.0, Y, lda_Y, SM_NONE, Y, lda_Y, SM_NONE, 0.0, G, b, part));
  } else {
    CQR_TRY(launch_fill2d(c, G, b, b, b, 0.0));
  }
  return launch_tril_halve(c, G, (int)b, T, lda_T);
}

int cqr_compute_invT_from_Y2(cqr_ctx* c, const double* Y1, int64_t lda_Y1, const double* Y2, int64_t lda_Y2,
                             double* T, int64_t lda_T, int64_t mb, int64_t b1, int64_t b2) {
  CQR_CHECK_CTX(c);
  if (!Y1 || !Y2 || !T || b1 < 1 || b2 < 1 || mb < 0 || lda_Y1 < mb || lda_Y2 < mb || lda_T < b1)
    return set_err(c, CQR_ERR_INVALID, "compute_invT_from_Y (two Y): bad arguments");
  if (mb == 0) {  // no rows: the product is zero
    return launch_fill2d(c, T, lda_T, b1, b2, 0.0);
  }
  const size_t sd = gemm_f64_scratch_doubles(b1, b2, mb);
  CQR_TRY(ensure_ws(c, 8 * (sd + 64)));
  return gemm_f64_ws(c, 1, 0, b1, b2, mb, 1.0, Y1, lda_Y1, SM_NONE, Y2, lda_Y2, SM_NONE, 0.0, T, lda_T,
                     sd ? static_cast<double*>(c->ws) : nullptr);
}

static int upd_core(cqr_ctx* c, const double* Y, int64_t lda_Y, int implicitY, double* A, int64_t lda_A,
                    int64_t mb, int64_t kb, int64_t b, const double* W, int W_is_T) {
  const int64_t bb = b * b;
  const size_t sd = std::max(gemm_f64_scratch_doubles(b, b, mb), wy_update_ws_doubles(mb, kb, (int)b));
  Sizer s;
  const size_t oG = s.add(8 * bb), oI = s.add(8 * bb), oT = s.add(8 * bb), oW = s.add(8 * bb),
               oS = s.add(8 * (sd + 64));
  CQR_TRY(ensure_ws(c, s.bytes));
  char* base = (char*)c->ws;
  auto at = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
  const int bi = (int)b;
  const int smY = implicitY ? SM_UNIT_LOWER : SM_NONE;
  const double* invTT;
  if (!W) {                                            // compute_invT_from_Y (qr_2d.cxx:242-246)
    CQR_TRY(gemm_f64_ws(c, 1, 0, b, b, mb, 1.0, Y, lda_Y, smY, Y, lda_Y, smY, 0.0, at(oG), b, at(oS)));
    CQR_TRY(launch_tril_halve(c, at(oG), bi, at(oI), b));
    invTT = at(oI);
  } else if (W_is_T) {                                 // W already holds T⁻ᵀ (qr_2d.cxx:248)
    invTT = W;
  } else {                                             // comp_bcast_T_from_W (qr_2d.cxx:179-208)
    CQR_TRY(invTT_from_W(c, W, bi, Y, lda_Y, smY, at(oW), at(oI)));
    invTT = at(oI);
  }
  CQR_TRY(Tt_from_invTT(c, invTT, b, at(oT), bi));
  return wy_update_f64_ws(c, Y, lda_Y, implicitY, at(oT), A, lda_A, mb, kb, bi, at(oS));
}

int cqr_upd_A(cqr_ctx* c, const double* Y, int64_t lda_Y, double* A, int64_t lda_A, int64_t mb, int64_t kb,
              int64_t b, const double* W, int W_is_T) {
  CQR_CHECK_CTX(c);
  if (!Y || !A || b < 1 || mb < 0 || kb < 0 || lda_Y < mb || lda_A < mb || mb < b)
    return set_err(c, CQR_ERR_INVALID, "upd_A: bad arguments");
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "upd_A: b > %d", CQR_MAX_B);
  if (kb == 0) return CQR_OK;
  return upd_core(c, Y, lda_Y, 0, A, lda_A, mb, kb, b, W, W_is_T);
}

int cqr_update_A(cqr_ctx* c, const double* Y, int64_t lda_Y, double* A, int64_t lda_A, int64_t m, int64_t k,
                 int64_t b, const double* W, double* aggreg_Y, int64_t lda_aY, int W_is_T) {
  CQR_CHECK_CTX(c);
  if (!Y || !A || b < 1 || m < b || k < 0 || lda_Y < m || lda_A < m || (aggreg_Y && lda_aY < m))
    return set_err(c, CQR_ERR_INVALID, "update_A: bad arguments");
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "update_A: b > %d", CQR_MAX_B);
  if (k > 0) CQR_TRY(upd_core(c, Y, lda_Y, 1, A, lda_A, m, k, b, W, W_is_T));
  if (aggreg_Y) CQR_TRY(launch_copy_unit_lower(c, Y, lda_Y, aggreg_Y, lda_aY, m, b));
  return CQR_OK;
}

}  // extern "C"

namespace cqr {

// Workspace of QR_2D for an m×k matrix with panel width b: the maximum over
// all panels (the split-K scratch of the update is not monotone in j).
// part 0: the whole trailing block; 1: the next panel's columns; 2: the rest.
static size_t max_update_ws(int64_t m, int64_t k, int64_t b, int part = 0) {
  size_t mx = 0;
  for (int64_t j = 0; j < k; j += b) {
    const int64_t mj = m - j, kj = k - j, bj = std::min(kj, b);
    if (!(kj - b > 0 && mj - b > 0)) continue;
    const int64_t bn = std::min(b, kj - bj), cols = part == 0 ? kj - bj : part == 1 ? bn : kj - bj - bn;
    if (cols > 0) mx = std::max(mx, wy_update_ws_doubles(mj, cols, (int)bj));
  }
  return mx;
}

static size_t max_panel_ws(int64_t m, int64_t k, int64_t b, int64_t leaf_rows) {
  size_t mx = 0;
  for (int64_t j = 0; j < k; j += b) mx = std::max(mx, panel_ws_bytes(m - j, std::min(k - j, b), 0, leaf_rows));
  return mx;
}

size_t qr2d_ws(int64_t m, int64_t k, int64_t b, int64_t leaf_rows, Qr2dWS* o) {
  Sizer s;
  const int64_t b0 = std::min(k, b);
  o->panel = s.add(max_panel_ws(m, k, b, leaf_rows));
  o->upd_near = s.add(8 * (std::max(max_update_ws(m, k, b, 0), max_update_ws(m, k, b, 1)) + 64));
  o->upd_far = s.add(8 * (max_update_ws(m, k, b, 2) + 64));
  o->W = s.add(8 * b0 * b0);
  o->Tt = s.add(8 * 2 * b0 * b0);
  o->Ye = s.add(8 * 4 * (size_t)m * b0);   // explicit Y and Y·Tt, two panels each
  return s.bytes;
}

// Issues the kernels of its scope on the context's side stream.
struct OnSide {
  cqr_ctx* c;
  hipStream_t main;
  explicit OnSide(cqr_ctx* ctx) : c(ctx), main(ctx->stream) { c->stream = c->side; }
  ~OnSide() { c->stream = main; }
};

int qr2d_run(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t k, int64_t b, double* aggreg_Y,
             int64_t lda_aY, char* base, const Qr2dWS& o, double* Tt_out, int prec) {
  double* W = reinterpret_cast<double*>(base + o.W);
  double* Tt2[2] = {reinterpret_cast<double*>(base + o.Tt), reinterpret_cast<double*>(base + o.Tt) + b * b};
  double* ws_near = reinterpret_cast<double*>(base + o.upd_near);
  double* ws_far = reinterpret_cast<double*>(base + o.upd_far);
  const bool la = c->lookahead && k > 2 * b && m > 2 * b;
  // Each panel's Y is made explicit (unit lower) and Y·Tt formed once (in
  // FP64), so both updates are two plain products: Z = Yᵀ·A (split-K) and
  // A −= (Y·Tt)·Z (the rank-b update kernel, or the mixed-precision kernels)
  const bool expl = true;
  double* Ye2[2] = {reinterpret_cast<double*>(base + o.Ye),
                    reinterpret_cast<double*>(base + o.Ye) + 2 * m * std::min(k, b)};
  if (la) {  // the side stream starts after everything already queued
    CQR_TRY(ensure_side(c));
    CQR_HIP(c, hipEventRecord(c->ev_fork, c->stream));
    CQR_HIP(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
  }
  int rc = CQR_OK;
  // The far update's second product (A −= (Y·Tᵀ)·Z, whose workgroups fill
  // every CU's registers) is held back until the next panel's construct_Q1
  // has finished, so the HR core and the small launches that follow find
  // free CUs at once; Z = Yᵀ·A still overlaps the leaves, tree and Q1.
  // (Measured at 32768×4096, b = 128: 43.7 ms against 45.6 ms ungated,
  // 48.2 ms when gated after the HR core.)
  struct Pending {
    bool on = false;
    const double* YT;
    int64_t m, n;
    int b;
    const double* Z;
    double* A;
  } pend;
  const int gate = la && expl ? 1 : 0;
  auto flush = [&](bool gated) -> int {
    if (!pend.on) return CQR_OK;
    pend.on = false;
    if (gated) CQR_HIP(c, hipStreamWaitEvent(c->side, c->ev_gate, 0));
    {
      OnSide g(c);
      CQR_TRY(wy_explicit_apply(c, pend.YT, pend.m, pend.n, pend.b, pend.Z, pend.A, lda, prec));
    }
    CQR_HIP(c, hipEventRecord(c->ev_join, c->side));
    return CQR_OK;
  };
  c->gate_at = gate;
  for (int64_t j = 0, p = 0;; j += b, ++p) {             // qr_2d.cxx:299-363
    const int64_t mj = m - j, kj = k - j, bj = std::min(kj, b);
    double* P = A + j + j * lda;
    double* Tt = Tt_out ? Tt_out + p * b * b : Tt2[p & 1];
    // the explicit Y and Y·Tt come out of the panel's own Y product
    // whenever a trailing update follows
    const bool more = kj - b > 0 && mj - b > 0;
    double* Ye = Ye2[p & 1];
    double* YT = Ye + mj * bj;
    if ((rc = hh_recon_panel(c, P, lda, mj, bj, W, nullptr, nullptr, Tt, 0, base + o.panel, nullptr,
                             more && expl ? Ye : nullptr, more && expl ? YT : nullptr)))
      break;
    if ((rc = flush(true))) break;
    if (aggreg_Y && (rc = launch_copy_unit_lower(c, P, lda, aggreg_Y + j + j * lda_aY, lda_aY, mj, bj))) break;
    if (!more) break;
    auto update = [&](int64_t col0, int64_t ncols, double* ws) -> int {
      if (expl) return wy_update_explicit(c, Ye, YT, mj, ncols, (int)bj, P + col0 * lda, lda, ws, prec);
      return wy_update_ws(c, P, lda, 1, Tt, 0, P + col0 * lda, lda, mj, ncols, (int)bj, ws, prec);
    };
    if (!la) {
      if ((rc = update(bj, kj - bj, ws_near))) break;
      continue;
    }
    // Lookahead (qr_2d.cxx:378-635): the next panel's b columns first, on
    // the main stream, once the far update of the previous panel (which
    // covered them) has finished; then the far columns on the side stream,
    // overlapping the next panel's factorization.  Tᵀ alternates between
    // two buffers because the side stream still reads the previous one.
    const int64_t bn = std::min(b, kj - bj), rest = kj - bj - bn;
    if (p > 0 && (rc = hipStreamWaitEvent(c->stream, c->ev_join, 0)) != hipSuccess) {
      rc = set_err(c, CQR_ERR_DEVICE, "QR_2D: stream wait failed");
      break;
    }
    if ((rc = update(bj, bn, ws_near))) break;
    if (rest > 0) {
      if (hipEventRecord(c->ev_fork, c->stream) != hipSuccess ||
          hipStreamWaitEvent(c->side, c->ev_fork, 0) != hipSuccess) {
        rc = set_err(c, CQR_ERR_DEVICE, "QR_2D: stream fork failed");
        break;
      }
      if (gate) {
        OnSide g(c);
        double* Z = nullptr;
        if ((rc = wy_explicit_yta(c, Ye, mj, rest, (int)bj, P + (bj + bn) * lda, lda, ws_far, prec, &Z))) break;
        pend.on = true;
        pend.YT = YT; pend.m = mj; pend.n = rest; pend.b = (int)bj; pend.Z = Z; pend.A = P + (bj + bn) * lda;
        continue;
      }
      OnSide g(c);
      if ((rc = update(bj + bn, rest, ws_far))) break;
      if (hipEventRecord(c->ev_join, c->side) != hipSuccess) {
        rc = set_err(c, CQR_ERR_DEVICE, "QR_2D: event record failed");
        break;
      }
    }
  }
  if (rc == CQR_OK) rc = flush(false);
  c->gate_at = 0;
  if (la) {  // join: work queued on the caller's stream after this sees all of it
    CQR_HIP(c, hipEventRecord(c->ev_join, c->side));
    CQR_HIP(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
  }
  return rc;
}

}  // namespace cqr

extern "C" {

int cqr_qr_2d(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t k, int64_t b, double* aggreg_Y,
              int64_t lda_aY) {
  CQR_CHECK_CTX(c);
  if (!A || m < 1 || k < 1 || b < 1 || lda < m || m < k || (aggreg_Y && lda_aY < m))
    return set_err(c, CQR_ERR_INVALID, "QR_2D: bad arguments (needs m >= k, lda >= m)");
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "QR_2D: b > %d", CQR_MAX_B);
  CQR_TRY(finite_check(c, A, lda, m, k));
  Qr2dWS o;
  const size_t need = qr2d_ws(m, k, b, c->leaf_rows, &o);
  CQR_TRY(ensure_ws(c, need));
  return qr2d_run(c, A, lda, m, k, b, aggreg_Y, lda_aY, (char*)c->ws, o, nullptr, CQR_PREC_F64);
}

int cqr_geqrf2d(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t k, int64_t b, double* Tt, int prec) {
  CQR_CHECK_CTX(c);
  if (!A || m < 1 || k < 1 || b < 1 || lda < m || m < k)
    return set_err(c, CQR_ERR_INVALID, "geqrf2d: bad arguments (needs m >= k, lda >= m)");
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "geqrf2d: b > %d", CQR_MAX_B);
  if (prec != CQR_PREC_F64 && prec != CQR_PREC_F32 && prec != CQR_PREC_BF16)
    return set_err(c, CQR_ERR_UNSUPPORTED, "geqrf2d: precision %d", prec);
  CQR_TRY(finite_check(c, A, lda, m, k));
  Qr2dWS o;
  const size_t need = qr2d_ws(m, k, b, c->leaf_rows, &o);
  CQR_TRY(ensure_ws(c, need));
  return qr2d_run(c, A, lda, m, k, b, nullptr, 0, (char*)c->ws, o, Tt, prec);
}

int cqr_qr_2d_2d(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t k, int64_t b, int64_t b_sub,
                 double* aggreg_Y, int64_t lda_aY) {
  CQR_CHECK_CTX(c);
  if (!A || m < 1 || k < 1 || b < 1 || b_sub < 1 || lda < m || m < k || (b % b_sub) != 0 ||
      (aggreg_Y && lda_aY < m))
    return set_err(c, CQR_ERR_INVALID, "QR_2D_2D: bad arguments (needs m >= k, b %% b_sub == 0)");
  if (b_sub > CQR_MAX_B || b > CQR_MAX_B)
    return set_err(c, CQR_ERR_UNSUPPORTED, "QR_2D_2D: b > %d (aggregated T is one b×b block)", CQR_MAX_B);
  CQR_TRY(finite_check(c, A, lda, m, k));
  Qr2dWS o;
  const size_t q = qr2d_ws(m, std::min(k, b), b_sub, c->leaf_rows, &o);
  Sizer s;
  s.bytes = q;
  const int64_t bb = b * b;
  const size_t oG = s.add(8 * bb), oI = s.add(8 * bb), oTt = s.add(8 * bb);
  const size_t oS = s.add(8 * (std::max(gemm_f64_scratch_doubles(b, b, m), max_update_ws(m, k, b)) + 64));
  CQR_TRY(ensure_ws(c, s.bytes));
  char* base = (char*)c->ws;
  auto at = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
  for (int64_t j = 0;; j += b) {                         // qr_2d.cxx:772-897
    const int64_t mj = m - j, kj = k - j;
    double* P = A + j + j * lda;
    double* aY = aggreg_Y ? aggreg_Y + j + j * lda_aY : nullptr;
    if (kj - b > 0 && mj - b > 0) {
      CQR_TRY(qr2d_run(c, P, lda, mj, b, b_sub, aY, lda_aY, base, o, nullptr, CQR_PREC_F64));
      // aggregated T⁻ᵀ = tril(YᵀY), diagonal halved (compute_invT_from_Y)
      CQR_TRY(gemm_f64_ws(c, 1, 0, b, b, mj, 1.0, P, lda, SM_UNIT_LOWER, P, lda, SM_UNIT_LOWER, 0.0, at(oG), b,
                          at(oS)));
      CQR_TRY(launch_tril_halve(c, at(oG), (int)b, at(oI), b));
      CQR_TRY(Tt_from_invTT(c, at(oI), b, at(oTt), (int)b));
      CQR_TRY(wy_update_f64_ws(c, P, lda, 1, at(oTt), P + b * lda, lda, mj, kj - b, (int)b, at(oS)));
    } else {
      return qr2d_run(c, P, lda, mj, kj, b_sub, aY, lda_aY, base, o, nullptr, CQR_PREC_F64);
    }
  }
}

int cqr_tau_recon(cqr_ctx* c, char structure, int64_t ncol, int64_t nrow, int64_t lda, const double* Y,
                  double* tau) {
  CQR_CHECK_CTX(c);
  if (!Y || !tau || ncol < 0 || nrow < 0 || lda < 1) return set_err(c, CQR_ERR_INVALID, "tau_recon: bad arguments");
  if (structure != 'L' && structure != 'U' && structure != 'R')
    return set_err(c, CQR_ERR_INVALID, "tau_recon: structure must be 'R', 'L' or 'U'");
  return launch_tau_recon(c, structure, ncol, nrow, lda, Y, tau);
}

int cqr_pack_upper(cqr_ctx* c, const double* R, double* P, int64_t b, int64_t lda) {
  CQR_CHECK_CTX(c);
  if (!R || !P || b < 0 || lda < b) return CQR_ERR_INVALID;
  return launch_pack(c, 0, R, P, b, lda, 1);
}
int cqr_unpack_upper(cqr_ctx* c, const double* P, double* R, int64_t b, int64_t lda) {
  CQR_CHECK_CTX(c);
  if (!R || !P || b < 0 || lda < b) return CQR_ERR_INVALID;
  return launch_pack(c, 1, P, R, b, lda, 1);
}
int cqr_pack_lower(cqr_ctx* c, const double* T, double* P, int64_t b, int64_t lda, int has_diag) {
  CQR_CHECK_CTX(c);
  if (!T || !P || b < 0 || lda < b) return CQR_ERR_INVALID;
  return launch_pack(c, 2, T, P, b, lda, has_diag ? 1 : 0);
}
int cqr_unpack_lower(cqr_ctx* c, const double* P, double* T, int64_t b, int64_t lda, int has_diag) {
  CQR_CHECK_CTX(c);
  if (!T || !P || b < 0 || lda < b) return CQR_ERR_INVALID;
  return launch_pack(c, 3, P, T, b, lda, has_diag ? 1 : 0);
}

}  // extern "C"
