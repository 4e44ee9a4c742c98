// The implicit TSQR Q of cqr_bitree_tsqr applied to a block of columns:
//   B := Qᵀ·B   apply_tsqr_QT (alg/QR/tsqr/apply_tsqr_QT.cxx:31-165)
//   B := Q·B    the tree walk + leaf expansion of construct_Q1
//               (alg/QR/tsqr/construct_tsqr_Q.cxx:146-233) from any start block,
// and the callers built on them: construct_Q1 for k ≠ b or with ID, the
// tree-updated 2D QR (alg/QR/qr_2d/qr_tree_2d.cxx:30-124), and the export of
// one virtual rank's tree factors in the reference's packed layout.
//
// Q = Q_leaves·Q_tree.  Q_leaves is block diagonal: leaf l contributes the
// compact-WY factor I − V_l·T_l·V_lᵀ on its rows (V_l unit lower, stored as
// the leaf's [V\R] block).  Q_tree is the product of the node transforms in
// level order; node (z, s) acts on the top b rows of leaves z and s as
// Q_node = I − [I; V2]·T·[I; V2]ᵀ with V2 and T upper triangular.
//   Qᵀ·B: the leaves (B_l −= V_l·T_lᵀ·(V_lᵀ·B_l)), then the levels forward:
//           X = Tᵀ·(B_z + V2ᵀ·B_s);  B_z −= X;  B_s −= V2·X
//   Q·B:  the levels backward with T in place of Tᵀ, then the leaves with T_l.
// Every step is one batched GEMM launch over all leaves or over one level's
// pairs (bgemm, addressing leaves by their row offsets); B is processed in
// column chunks that bound the scratch to 2·L·b·kc doubles.
#include "cqr_device.h"
#include "cqr_internal.h"
#include "leaf_body.h"

#include <vector>

namespace cqr {

namespace {

// A(0:b, 0:b) upper triangle := R's (copy_upper of qr_tree_2d.cxx:67-68).
__global__ void copy_upper_kernel(const double* __restrict__ R, int64_t ldr, double* __restrict__ A, int64_t lda,
                                  int b) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= b * b) return;
  const int r = e % b, c = e / b;
  if (r <= c) A[r + (int64_t)c * lda] = R[r + (int64_t)c * ldr];
}

// R (b×b, ld ldr) := the upper triangle of A's top b×b block, zeros below
// (pack_upper + unpack_upper into a zero buffer, bitree_tsqr.cxx:143-145).
__global__ void upper_zero_kernel(const double* __restrict__ A, int64_t lda, double* __restrict__ R, int64_t ldr,
                                  int b) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= b * b) return;
  const int r = e % b, c = e / b;
  R[r + (int64_t)c * ldr] = r <= c ? A[r + (int64_t)c * lda] : 0.0;
}

// Tt := Tᵀ of the compact-WY factor of b reflectors from their Gram matrix
// G = VᵀV (b×b, ld b) and tau (LAPACK dlarft('F','C')): T(i,i) = τ_i and
// T(0:i, i) = −τ_i·T(0:i, 0:i)·G(0:i, i), column by column.  One workgroup;
// thread r owns row r of T, kept in LDS (b·b doubles).  A zero τ leaves its
// row and column of T zero (the reflector is the identity).
__global__ void __launch_bounds__(512) larft_kernel(const double* __restrict__ G, const double* __restrict__ tau,
                                                    double* __restrict__ Tt, int b) {
  extern __shared__ double Ts[];
  const int r = threadIdx.x;
  for (int e = r; e < b * b; e += blockDim.x) Ts[e] = 0.0;
  __syncthreads();
  for (int i = 0; i < b; ++i) {
    const double ti = tau[i];
    if (r < i) {
      double s = 0.0;
      for (int k = r; k < i; ++k) s += Ts[r + k * b] * G[k + (int64_t)i * b];
      Ts[r + i * b] = -ti * s;
    } else if (r == i) {
      Ts[i + i * b] = ti;
    }
    __syncthreads();
  }
  for (int e = r; e < b * b; e += blockDim.x) {
    const int rr = e % b, cc = e / b;
    Tt[cc + rr * b] = Ts[e];
  }
}

// Q (m×k) := [ID; 0] (ID b×k, ld ldid) or the first k columns of I_m.
__global__ void start_block_kernel(double* __restrict__ Q, int64_t ldq, int64_t m, int64_t k,
                                   const double* __restrict__ ID, int64_t ldid, int b) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * k;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % m, j = e / m;
    double v;
    if (ID) v = i < b ? ID[i + j * ldid] : 0.0;
    else v = i == j ? 1.0 : 0.0;
    Q[i + j * ldq] = v;
  }
}

BgemmArgs bga(int m, int n, int k, double alpha, double beta, const TsqrPlan& p) {
  BgemmArgs g{};
  g.m = m; g.n = n; g.k = k; g.alpha = alpha; g.beta = beta;
  g.pm = p.m; g.pL = (int)p.nleaves;
  return g;
}

// column chunk: the scratch 2·L·b·kc doubles stays within 256 MiB
int64_t chunk_cols(const TsqrPlan& p, int64_t k) {
  const int64_t per_col = 2 * p.nleaves * p.b * 8;
  int64_t kc = ((int64_t)256 << 20) / std::max<int64_t>(per_col, 1);
  kc = std::max<int64_t>(64, kc / 64 * 64);
  return std::min(kc, std::max<int64_t>(k, 1));
}

size_t apply_scratch_bytes(const TsqrPlan& p, int64_t k) {
  return (size_t)(2 * p.nleaves * p.b * chunk_cols(p, k)) * 8;
}

// leaves: B_l := (I − V_l·op(T_l)·V_lᵀ)·B_l, op = ᵀ for Qᵀ
int apply_leaves(cqr_ctx* c, int trans, const TsqrPlan& p, const double* Y, int64_t ldy, const double* leafT,
                 double* B, int64_t ldb, int n, double* W, double* W2) {
  const int b = (int)p.b, L = (int)p.nleaves;
  const int64_t bb = (int64_t)b * b, wz = (int64_t)b * n;
  {  // W_l = V_lᵀ·B_l (contraction over the leaf's rows)
    BgemmArgs g = bga(b, n, -1, 1.0, 0.0, p);
    g.A = bmat(Y, ldy, 0, BM_LEAFROWS); g.ta = 1; g.smA = SM_UNIT_LOWER;
    g.B = bmat(B, ldb, 0, BM_LEAFROWS);
    g.C = bmat(W, b, wz);
    CQR_TRY(bgemm(c, g, L, b));
  }
  {  // W2_l = op(T_l)·W_l
    BgemmArgs g = bga(b, n, b, 1.0, 0.0, p);
    g.A = bmat(leafT, b, bb); g.ta = trans; g.smA = SM_UPPER;
    g.B = bmat(W, b, wz);
    g.C = bmat(W2, b, wz);
    CQR_TRY(bgemm(c, g, L, b));
  }
  // B_l −= V_l·W2_l
  BgemmArgs g = bga(-1, n, b, -1.0, 1.0, p);
  g.A = bmat(Y, ldy, 0, BM_LEAFROWS); g.smA = SM_UNIT_LOWER;
  g.B = bmat(W2, b, wz);
  g.C = bmat(B, ldb, 0, BM_LEAFROWS);
  return bgemm(c, g, L, (int)p.max_leaf_rows);
}

// one tree level: every pair (z, s = z + ⌈n/2⌉) of leaves' top b rows
int apply_level(cqr_ctx* c, int trans, const TsqrPlan& p, int lv, const double* nodeV2, const double* nodeT,
                double* B, int64_t ldb, int n, double* X, double* X2) {
  const int b = (int)p.b;
  const int npairs = (int)p.levels[lv].size();
  if (npairs == 0) return CQR_OK;
  const int half = p.levels[lv][0].second;  // sender of pair 0 = ⌈n/2⌉
  const int64_t bb = (int64_t)b * b, xz = (int64_t)b * n;
  const double* V2 = nodeV2 + p.level_node0[lv] * bb;
  const double* T = nodeT + p.level_node0[lv] * bb;
  {  // X = B_z
    BgemmArgs g = bga(b, n, b, 1.0, 0.0, p);
    g.A = bmat(V2, b, bb); g.smA = SM_IDENTITY;
    g.B = bmat(B, ldb, 0, BM_LEAFROWS);
    g.C = bmat(X, b, xz);
    CQR_TRY(bgemm(c, g, npairs, b));
  }
  {  // X += V2ᵀ·B_s
    BgemmArgs g = bga(b, n, b, 1.0, 1.0, p);
    g.A = bmat(V2, b, bb); g.ta = 1; g.smA = SM_UPPER;
    g.B = bmat(B, ldb, 0, BM_LEAFROWS_SHIFTED); g.zoff = half;
    g.C = bmat(X, b, xz);
    CQR_TRY(bgemm(c, g, npairs, b));
  }
  {  // X2 = op(T)·X
    BgemmArgs g = bga(b, n, b, 1.0, 0.0, p);
    g.A = bmat(T, b, bb); g.ta = trans; g.smA = SM_UPPER;
    g.B = bmat(X, b, xz);
    g.C = bmat(X2, b, xz);
    CQR_TRY(bgemm(c, g, npairs, b));
  }
  {  // B_z −= X2
    BgemmArgs g = bga(b, n, b, -1.0, 1.0, p);
    g.A = bmat(V2, b, bb); g.smA = SM_IDENTITY;
    g.B = bmat(X2, b, xz);
    g.C = bmat(B, ldb, 0, BM_LEAFROWS);
    CQR_TRY(bgemm(c, g, npairs, b));
  }
  // B_s −= V2·X2
  BgemmArgs g = bga(b, n, b, -1.0, 1.0, p);
  g.A = bmat(V2, b, bb); g.smA = SM_UPPER;
  g.B = bmat(X2, b, xz);
  g.C = bmat(B, ldb, 0, BM_LEAFROWS_SHIFTED); g.zoff = half;
  return bgemm(c, g, npairs, b);
}

// B (p.m × k) := Qᵀ·B (trans = 1) or Q·B (trans = 0); scratch of
// apply_scratch_bytes(p, k) bytes.
int tsqr_apply(cqr_ctx* c, int trans, const TsqrPlan& p, const double* Y, int64_t ldy, const double* leafT,
               const double* nodeV2, const double* nodeT, double* B, int64_t ldb, int64_t k, double* scratch) {
  const int64_t kc = chunk_cols(p, k);
  const int nlv = (int)p.levels.size();
  double* W = scratch;
  double* W2 = scratch + p.nleaves * p.b * kc;
  for (int64_t j0 = 0; j0 < k; j0 += kc) {
    const int n = (int)std::min(kc, k - j0);
    double* Bj = B + j0 * ldb;
    if (trans) {
      CQR_TRY(apply_leaves(c, 1, p, Y, ldy, leafT, Bj, ldb, n, W, W2));
      for (int lv = 0; lv < nlv; ++lv) CQR_TRY(apply_level(c, 1, p, lv, nodeV2, nodeT, Bj, ldb, n, W, W2));
    } else {
      for (int lv = nlv - 1; lv >= 0; --lv) CQR_TRY(apply_level(c, 0, p, lv, nodeV2, nodeT, Bj, ldb, n, W, W2));
      CQR_TRY(apply_leaves(c, 0, p, Y, ldy, leafT, Bj, ldb, n, W, W2));
    }
  }
  return CQR_OK;
}

// One rank (np_work == 1): Y (m×b, ld ldy) is the block's dgeqrf [V\R] and
// tau its b taus, so Q = H_1···H_b = I − V·T·Vᵀ with T = dlarft(V, tau).
// B (m×k) := Qᵀ·B (trans = 1, dormqr('L','T'), apply_tsqr_QT.cxx:73-80) or
// Q·B (trans = 0, the dorgqr / dormqr('L','N') of construct_tsqr_Q.cxx:87-102).
size_t local_q_ws_bytes(int64_t m, int64_t b, int64_t k) {
  Sizer s;
  s.add(8 * b * b);
  s.add(8 * b * b);
  s.add(8 * (gemm_f64_scratch_doubles(b, b, m) + 64));
  s.add(8 * (wy_update_ws_doubles(m, k, (int)b) + 64));
  return s.bytes;
}

int local_q_apply(cqr_ctx* c, int trans, const double* Y, int64_t ldy, const double* tau, double* B, int64_t ldb,
                  int64_t m, int64_t b, int64_t k, char* base) {
  Sizer s;
  const size_t oG = s.add(8 * b * b), oT = s.add(8 * b * b), oP = s.add(8 * (gemm_f64_scratch_doubles(b, b, m) + 64)),
               oW = s.add(8 * (wy_update_ws_doubles(m, k, (int)b) + 64));
  auto at = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
  CQR_TRY(gemm_f64_ws(c, 1, 0, b, b, m, 1.0, Y, ldy, SM_UNIT_LOWER, Y, ldy, SM_UNIT_LOWER, 0.0, at(oG), b, at(oP)));
  const int bi = (int)b;
  CQR_TRY(set_max_lds(c, (const void*)larft_kernel, 163840));
  hipLaunchKernelGGL(larft_kernel, dim3(1), dim3(std::max(64, std::min(512, (bi + 63) / 64 * 64))),
                     sizeof(double) * (size_t)b * b, c->stream, at(oG), tau, at(oT), bi);
  CQR_TRY(check_launch(c, "larft"));
  return wy_update_ws(c, Y, ldy, 1, at(oT), trans ? 0 : 1, B, ldb, m, k, bi, at(oW), CQR_PREC_F64);
}

}  // namespace

// bitree_tsqr without tree_data (bitree_tsqr.cxx:91-248) over P virtual
// ranks.  Rank r holds rows [off_r, off_r + mb_r) with mb_r from the
// reference's distribution (:110-111, root 0); rows past the last full b
// block go to the last working rank.  One working rank: dgeqrf of the whole
// block (:129-146; A keeps its input unless output_Y).  Several: every rank's
// block gets its dgeqrf [V\R] in place (:156); the tree over the ranks' R
// (pairs (r, r + ⌈n/2⌉), :172-232) runs as one launch per level, and with
// output_Y each level writes the node's R into the receiver's top triangle
// and the node's reflector block V2 into the sender's (:184-189, :223-229).
// tau receives rank 0's b taus; R the final R, zero below the diagonal.
int bitree_ranks(cqr_ctx* c, double* A, int64_t lda, double* R, double* tau, int64_t m, int64_t b, int64_t P,
                 int output_Y) {
  const int64_t nw = ranks_working(m, b, P), bb = b * b;
  std::vector<int64_t> off(nw + 1, 0);
  for (int64_t r = 0; r < nw; ++r) {
    int64_t mb = ((m / b) / P) * b;
    if ((m / b) % P > r) mb += b;
    off[r + 1] = off[r] + mb;
  }
  off[nw] = m;  // the rows past the last full block
  int64_t maxrows = 0;
  for (int64_t r = 0; r < nw; ++r) maxrows = std::max(maxrows, off[r + 1] - off[r]);
  Sizer s;
  const size_t oL = s.add(local_tsqr_ws_bytes(maxrows, b, c->leaf_rows));
  const size_t oR = s.add(8 * nw * bb), oV = s.add(8 * nw * bb), oT = s.add(8 * nw * bb), oNt = s.add(8 * nw * b),
               oTau = s.add(8 * b), oC = s.add(nw == 1 && !output_Y ? 8 * (size_t)m * b : 0);
  CQR_TRY(ensure_ws(c, s.bytes));
  char* base = (char*)c->ws;
  auto at = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
  const int bi = (int)b;
  const unsigned gb = (unsigned)((bb + 255) / 256);
  if (nw == 1) {
    double* W = A;
    int64_t ldw = lda;
    if (!output_Y) {  // the reference factors a copy (R_buf)
      W = at(oC);
      ldw = m;
      CQR_TRY(launch_copy2d(c, A, lda, W, m, m, b));
    }
    CQR_TRY(local_tsqr_ws(c, W, ldw, m, b, tau, nullptr, base + oL));
    hipLaunchKernelGGL(upper_zero_kernel, dim3(gb), dim3(256), 0, c->stream, W, ldw, R, b, bi);
    return check_launch(c, "upper_zero");
  }
  double* Rl = at(oR);
  for (int64_t r = 0; r < nw; ++r) {
    double* Ar = A + off[r];
    CQR_TRY(local_tsqr_ws(c, Ar, lda, off[r + 1] - off[r], b, r == 0 ? tau : at(oTau), nullptr, base + oL));
    hipLaunchKernelGGL(upper_zero_kernel, dim3(gb), dim3(256), 0, c->stream, Ar, lda, Rl + r * bb, b, bi);
    CQR_TRY(check_launch(c, "upper_zero"));
  }
  int node = 0;
  for (int64_t n = nw; n > 1; n = (n + 1) / 2) {
    const int64_t half = (n + 1) / 2, npairs = n / 2;
    CQR_TRY(tree_level(c, Rl, (int)n, node, bi, at(oV), at(oT), at(oNt)));
    if (output_Y) {
      for (int64_t z = 0; z < npairs; ++z) {
        hipLaunchKernelGGL(copy_upper_kernel, dim3(gb), dim3(256), 0, c->stream, Rl + z * bb, b, A + off[z], lda, bi);
        hipLaunchKernelGGL(copy_upper_kernel, dim3(gb), dim3(256), 0, c->stream, at(oV) + (node + z) * bb, b,
                           A + off[z + half], lda, bi);
      }
      CQR_TRY(check_launch(c, "copy_upper"));
    }
    node += (int)npairs;
  }
  hipLaunchKernelGGL(upper_zero_kernel, dim3(gb), dim3(256), 0, c->stream, Rl, b, R, b, bi);
  return check_launch(c, "upper_zero");
}

namespace {

struct TreeView {
  const double *leafT, *nodeV2, *nodeT;
};
TreeView tree_view(const double* tree_data, const TsqrPlan& p) {
  const int64_t bb = p.b * p.b;
  TreeView v;
  v.leafT = tree_data;
  v.nodeV2 = v.leafT + p.nleaves * bb;
  v.nodeT = v.nodeV2 + p.nnodes * bb;
  return v;
}

int apply_entry(cqr_ctx* c, const char* fn, int trans, const double* Y, int64_t lda_Y, const double* tau, double* B,
                int64_t lda_B, int64_t m, int64_t b, int64_t k, int64_t nleaves, const double* tree_data) {
  CQR_CHECK_CTX(c);
  if (!Y || !B || b < 1 || m < b || lda_Y < m || lda_B < m || k < 0 || nleaves < 0)
    return set_err(c, CQR_ERR_INVALID, "%s: bad arguments (m=%lld b=%lld k=%lld)", fn, (long long)m,
                   (long long)b, (long long)k);
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "%s: b=%lld > %d", fn, (long long)b, CQR_MAX_B);
  if (!tree_data) {  // the one-rank branch: Y, tau of one dgeqrf
    if (ranks_working(m, b, std::max<int64_t>(nleaves, 1)) != 1 || !tau)
      return set_err(c, CQR_ERR_INVALID, "%s: without tree_data only one rank (nleaves <= 1) and tau are allowed", fn);
    if (k == 0) return CQR_OK;
    CQR_TRY(ensure_ws(c, local_q_ws_bytes(m, b, k)));
    return local_q_apply(c, trans, Y, lda_Y, tau, B, lda_B, m, b, k, (char*)c->ws);
  }
  if (k == 0) return CQR_OK;
  TsqrPlan p;
  const int s = make_plan(m, b, nleaves, 0, p);
  if (s != CQR_OK) return set_err(c, s, "%s: cannot plan", fn);
  CQR_TRY(ensure_ws(c, apply_scratch_bytes(p, k)));
  const TreeView v = tree_view(tree_data, p);
  return tsqr_apply(c, trans, p, Y, lda_Y, v.leafT, v.nodeV2, v.nodeT, B, lda_B, k, static_cast<double*>(c->ws));
}

// One virtual rank: bitree_tsqr is dgeqrf of the panel and apply_tsqr_QT
// is dormqr('L','T') of the trailing columns (the np_work == 1 branches,
// bitree_tsqr.cxx:128-140 and apply_tsqr_QT.cxx:73-80).  A panel that fits
// one leaf is factored by the leaf kernel (T upper); a taller one by TSQR
// + Householder reconstruction, whose [Y\R] is dgeqrf's (Tᵀ).
int qr_tree_2d_one_rank(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t k, int64_t b, int64_t stop) {
  size_t nP = 0, nU = 0;
  for (int64_t j = 0; j < stop; j += b) {
    const int64_t mj = m - j, bw = std::min(b, stop - j), kt = k - j - bw;
    if (mj > leaf_capacity_rows(bw)) {
      const size_t pb = panel_ws_bytes(mj, bw, 0, c->leaf_rows);
      if (!pb) return set_err(c, CQR_ERR_UNSUPPORTED, "qr_tree_2d: no leaf plan for %lld×%lld", (long long)mj,
                              (long long)bw);
      nP = std::max(nP, pb);
    }
    if (kt > 0) nU = std::max(nU, wy_update_ws_doubles(mj, kt, (int)bw) * 8);
  }
  Sizer sz;
  const size_t oT = sz.add(8 * b * b), oTt = sz.add(8 * b * b), oR = sz.add(8 * b * b), oTau = sz.add(8 * b), oP = sz.add(nP),
               oU = sz.add(nU);
  CQR_TRY(ensure_ws(c, sz.bytes));
  char* base = (char*)c->ws;
  auto at = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
  for (int64_t j = 0; j < stop; j += b) {
    const int64_t mj = m - j, bw = std::min(b, stop - j), kt = k - j - bw;
    double* Aj = A + j + j * lda;
    if (mj <= leaf_capacity_rows(bw)) {
      CQR_TRY(leaf_qr(c, Aj, lda, mj, 1, (int)bw, at(oR), at(oTt), at(oTau)));
      // Tt := Tᵀ as I·Tᵀ
      BgemmArgs g{};
      g.m = g.n = g.k = (int)bw; g.alpha = 1.0; g.beta = 0.0;
      g.A = bmat(at(oTt), bw); g.smA = SM_IDENTITY;
      g.B = bmat(at(oTt), bw); g.tb = 1; g.smB = SM_UPPER;
      g.C = bmat(at(oT), bw);
      CQR_TRY(bgemm(c, g, 1, (int)bw));
    } else {
      CQR_TRY(hh_recon_panel(c, Aj, lda, mj, bw, at(oR), nullptr, nullptr, at(oT), 0, base + oP, nullptr));
    }
    if (kt > 0)  // at(oT) holds Tᵀ
      CQR_TRY(wy_update_ws(c, Aj, lda, 1, at(oT), 0, Aj + bw * lda, lda, mj, kt, (int)bw, at(oU), CQR_PREC_F64));
  }
  return CQR_OK;
}

}  // namespace

}  // namespace cqr

using namespace cqr;

extern "C" {

int cqr_apply_tsqr_QT(cqr_ctx* c, const double* Y, int64_t lda_Y, const double* tau, double* B, int64_t lda_B,
                      int64_t m, int64_t b, int64_t k, int64_t nleaves, const double* tree_data) {
  // with tree_data the leaves' compact-WY T carries their tau
  return apply_entry(c, "apply_tsqr_QT", 1, Y, lda_Y, tau, B, lda_B, m, b, k, nleaves, tree_data);
}

int cqr_apply_tsqr_Q(cqr_ctx* c, const double* Y, int64_t lda_Y, const double* tau, double* B, int64_t lda_B,
                     int64_t m, int64_t b, int64_t k, int64_t nleaves, const double* tree_data) {
  return apply_entry(c, "apply_tsqr_Q", 0, Y, lda_Y, tau, B, lda_B, m, b, k, nleaves, tree_data);
}

int cqr_construct_Q1_id(cqr_ctx* c, const double* Y, int64_t lda_Y, const double* tau, double* Q1,
                        int64_t lda_Q, int64_t m, int64_t b, int64_t k, int64_t nleaves, const double* tree_data,
                        const double* ID, int64_t ld_ID) {
  CQR_CHECK_CTX(c);
  // every argument is checked before the start block is written into Q1
  if (!Y || !Q1 || b < 1 || m < b || k < 1 || k > m || lda_Y < m || lda_Q < m || nleaves < 0 ||
      (ID && ld_ID < b) || (!tree_data && (!tau || ranks_working(m, b, std::max<int64_t>(nleaves, 1)) != 1)))
    return set_err(c, CQR_ERR_INVALID, "construct_Q1: bad arguments (m=%lld b=%lld k=%lld)", (long long)m,
                   (long long)b, (long long)k);
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "construct_Q1: b=%lld > %d", (long long)b, CQR_MAX_B);
  if (!ID && k == b && tree_data) return cqr_construct_Q1(c, Y, lda_Y, tau, Q1, lda_Q, m, b, k, nleaves, tree_data);
  const int gx = (int)std::min<int64_t>(4096, (m * k + 255) / 256);
  hipLaunchKernelGGL(start_block_kernel, dim3(gx), dim3(256), 0, c->stream, Q1, lda_Q, m, k, ID, ld_ID, (int)b);
  CQR_TRY(check_launch(c, "construct_Q1 start block"));
  return apply_entry(c, "construct_Q1", 0, Y, lda_Y, tau, Q1, lda_Q, m, b, k, nleaves, tree_data);
}

int cqr_qr_tree_2d(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t k, int64_t b, int64_t stop_at,
                   int64_t nleaves) {
  CQR_CHECK_CTX(c);
  if (!A || b < 1 || m < 1 || k < 1 || lda < m || nleaves < 0)
    return set_err(c, CQR_ERR_INVALID, "qr_tree_2d: bad arguments (m=%lld k=%lld b=%lld lda=%lld)",
                   (long long)m, (long long)k, (long long)b, (long long)lda);
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "qr_tree_2d: b=%lld > %d", (long long)b, CQR_MAX_B);
  const int64_t stop = stop_at >= 0 ? std::min(std::min(m, k), stop_at) : std::min(m, k);
  if (stop <= 0) return CQR_OK;
  CQR_TRY(finite_check(c, A, lda, m, k));
  if (nleaves == 1) return qr_tree_2d_one_rank(c, A, lda, m, k, b, stop);
  // one arena sized for the largest panel of each kind
  std::vector<TsqrPlan> plans;
  int64_t nLR = 0, nTau = 0, nNR = 0, nNTau = 0;
  size_t nS = 0;
  for (int64_t j = 0; j < stop; j += b) {
    const int64_t mj = m - j, bw = std::min(b, stop - j), kt = k - j - bw;
    TsqrPlan p;
    const int s = make_plan(mj, bw, nleaves, 0, p);
    if (s != CQR_OK)
      return set_err(c, s, "qr_tree_2d: cannot plan the %lld×%lld panel", (long long)mj, (long long)bw);
    nLR = std::max(nLR, p.nleaves * bw * bw);
    nTau = std::max(nTau, p.nleaves * bw);
    nNR = std::max(nNR, std::max<int64_t>(p.nnodes, 1) * bw * bw);
    nNTau = std::max(nNTau, std::max<int64_t>(p.nnodes, 1) * bw);
    if (kt > 0) nS = std::max(nS, apply_scratch_bytes(p, kt));
    plans.push_back(std::move(p));
  }
  Sizer sz;
  const size_t oLR = sz.add(8 * nLR), oLT = sz.add(8 * nLR), oTau = sz.add(8 * nTau), oNR = sz.add(8 * nNR),
               oV2 = sz.add(8 * nNR), oNT = sz.add(8 * nNR), oNTau = sz.add(8 * nNTau), oS = sz.add(nS);
  CQR_TRY(ensure_ws(c, sz.bytes));
  char* base = (char*)c->ws;
  auto at = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
  for (size_t pi = 0; pi < plans.size(); ++pi) {
    const TsqrPlan& p = plans[pi];
    const int64_t j = (int64_t)pi * b, mj = m - j, bw = p.b, kt = k - j - bw;
    double* Aj = A + j + j * lda;
    const int L = (int)p.nleaves, nn = (int)p.nnodes;
    CQR_TRY(leaf_tree(c, Aj, lda, mj, L, (int)bw, at(oLR), at(oLT), at(oTau), at(oNR), at(oV2), at(oNT), at(oNTau),
                      nn));
    const double* R = nn > 0 ? at(oNR) + (int64_t)(nn - 1) * bw * bw : at(oLR);
    hipLaunchKernelGGL(copy_upper_kernel, dim3((unsigned)((bw * bw + 255) / 256)), dim3(256), 0, c->stream, R, bw,
                       Aj, lda, (int)bw);
    CQR_TRY(check_launch(c, "copy_upper"));
    if (kt > 0)
      CQR_TRY(tsqr_apply(c, 1, p, Aj, lda, at(oLT), at(oV2), at(oNT), Aj + bw * lda, lda, kt, at(oS)));
  }
  return CQR_OK;
}

int cqr_tree_data_export(cqr_ctx* c, const double* tree_data, int64_t m, int64_t b, int64_t nleaves, int64_t rank,
                         double* out, int64_t* ndoubles) {
  CQR_CHECK_CTX(c);
  TsqrPlan p;
  const int s = make_plan(m, b, nleaves, 0, p);
  if (s != CQR_OK) return set_err(c, s, "tree_data_export: cannot plan m=%lld b=%lld", (long long)m, (long long)b);
  if (rank < 0 || rank >= p.nleaves)
    return set_err(c, CQR_ERR_INVALID, "tree_data_export: rank %lld outside [0, %lld)", (long long)rank,
                   (long long)p.nleaves);
  const int64_t nb = std::min<int64_t>(b, 16), tsz = nb * b, vsz = b * (b + 1) / 2, bb = b * b;
  int64_t total = 0;
  for (const auto& lv : p.levels)
    if (rank < (int64_t)lv.size()) total += tsz + vsz;
  if (ndoubles) *ndoubles = total;
  if (!out) return CQR_OK;
  if (!tree_data) return set_err(c, CQR_ERR_INVALID, "tree_data_export: tree_data is NULL");
  const TreeView v = tree_view(tree_data, p);
  double* o = out;
  for (size_t lv = 0; lv < p.levels.size(); ++lv) {
    if (rank >= (int64_t)p.levels[lv].size()) continue;
    const int64_t node = p.level_node0[lv] + rank;
    CQR_TRY(launch_blocked_T(c, v.nodeT + node * bb, o, (int)b, (int)nb));
    CQR_TRY(launch_pack(c, 0, v.nodeV2 + node * bb, o + tsz, b, b, 1));
    o += tsz + vsz;
  }
  return CQR_OK;
}

}  // extern "C"
