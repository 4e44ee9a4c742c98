// Internal declarations shared by the host drivers and the HIP kernels of
// the CA-QR library.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/candmc_qr.h"

enum { PK_GEMM_YTA = 0, PK_GEMM_UPD = 1, PK_GEMM_SMALL = 2, PK_REDUCE = 3, PK_LEAF = 4,
       PK_TREE = 5, PK_HR = 6, PK_BGEMM = 7, PK_NKINDS = 8 };

struct ProfRec {
  hipEvent_t a, b;
  int kind;
  double flops, bytes;
};

struct cqr_ctx {
  int prof = 0;
  std::vector<ProfRec> recs;      // pending event pairs
  std::vector<hipEvent_t> pool;   // free events
  double acc_ms[PK_NKINDS] = {}, acc_flops[PK_NKINDS] = {}, acc_bytes[PK_NKINDS] = {};
  int64_t acc_n[PK_NKINDS] = {};
  int device = 0;
  int num_cu = 256;             // compute units of the device
  hipStream_t stream = nullptr;
  void* ws = nullptr;           // device workspace arena
  size_t ws_bytes = 0;
  int check_finite = 0;
  int64_t leaf_rows = 0;        // 0 = automatic
  int* d_flag = nullptr;        // device flag for non-finite scans / tree timeouts
  uint64_t* d_cnt = nullptr;    // column counters: tree nodes, then (second half) leaves
  int* d_terr = nullptr;        // set by a tree node whose wait for a child timed out
  uint64_t* d_qflags = nullptr; // per (node, column block) flags of construct_Q1
  uint64_t* d_sflags = nullptr; // per 128-row block flags of the one-launch triangular solve
  double* h_pin = nullptr;      // 64 doubles of pinned host memory for small device-to-host reads
  hipEvent_t ev_host = nullptr; // recorded after a copy into h_pin
  uint64_t q1_epoch = 0;
  uint64_t tree_epoch = 0;      // launch epoch folded into the counters
  hipStream_t side = nullptr;   // second stream of the QR_2D lookahead
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_gate = nullptr;
  int gate_at = 0;              // QR_2D lookahead: each panel records ev_gate after construct_Q1
  int lookahead = 1;
  void* noalloc = nullptr;      // this device's fixed scratch (noalloc.hip)
  int lsq_escalate = 1;         // refactor a stalled mixed solve in FP64
  int lsq_last_prec = -1;       // factorization precision of the last solve's X
  std::string err;
};

namespace cqr {
// C (M×N) := beta·C + alpha·P·Z for K ≤ 256, beta ∈ {0, 1}: P column-major
// (ld ldp), Z column-major K×N (ld ldz), on FP64 MFMA (gemm_update.hip).
// shape 0: 128×128 tiles, 1: 256×128, 2: 128×256.
int update_gemm(cqr_ctx* c, int64_t M, int64_t N, int64_t K, double alpha, const double* P, int64_t ldp,
                const double* Z, int64_t ldz, double beta, double* C, int64_t ldc, int kind, int shape);
// P·Z with Z = [Z1 | Z2] split at column n1: C1 := P·Z1 (and a copy in C1b
// when non-null), C2 := P·Z2, in one pass over P.
int update_gemm_split(cqr_ctx* c, int64_t M, int64_t N, int64_t n1, int64_t K, const double* P, int64_t ldp,
                      const double* Z, int64_t ldz, double* C1, int64_t ldc1, double* C1b, int64_t ldc1b, double* C2,
                      int64_t ldc2, int kind);

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
int set_err(cqr_ctx* c, int code, const char* fmt, ...);
#define CQR_HIP(ctx, call)                                                    \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::cqr::set_err(ctx, CQR_ERR_DEVICE, "%s: %s", #call,             \
                            hipGetErrorString(e_));                           \
  } while (0)
#define CQR_TRY(expr)                                                         \
  do {                                                                        \
    int s_ = (expr);                                                          \
    if (s_ != CQR_OK) return s_;                                              \
  } while (0)
int check_launch(cqr_ctx* c, const char* what);
// Raises the dynamic-LDS limit of kernel `fn` to `bytes` on the context's
// device, once per (kernel, device, bytes); safe from several host threads.
int set_max_lds(cqr_ctx* c, const void* fn, int bytes);

// Event bracketing of a launch when profiling is on (no-ops otherwise).
hipEvent_t prof_begin(cqr_ctx* c);
void prof_end(cqr_ctx* c, hipEvent_t a, int kind, double flops, double bytes);

// ---------------------------------------------------------------------------
// Workspace: a bump allocator over the context's arena.  Callers first
// total their needs with a Sizer, grow the arena once, then carve.
// ---------------------------------------------------------------------------
struct Sizer {
  size_t bytes = 0;
  size_t add(size_t n) {            // returns offset
    size_t off = (bytes + 255) & ~size_t(255);
    bytes = off + n;
    return off;
  }
};
// Admits a launch whose workgroups wait on each other (context.cpp): it
// may run beside other such launches of the device only while all of
// their blocks fit on it at once, else its stream first waits for the
// oldest running ones.  Construct one right before the launch, on the
// launching stream, with the launch's kernel, blocks, threads per block and
// dynamic LDS; it records the launch's completion when it goes out of
// scope.  status() != CQR_OK: an ordering wait could not be queued.
int spin_gate_init(cqr_ctx* c);
class SpinScope {
 public:
  SpinScope(cqr_ctx* c, const void* fn, int64_t blocks, int threads, size_t lds);
  ~SpinScope();
  SpinScope(const SpinScope&) = delete;
  SpinScope& operator=(const SpinScope&) = delete;
  int status() const { return status_; }

 private:
  cqr_ctx* c;
  bool held = false;
  int need = 0;
  int status_ = CQR_OK;
};
// the lookahead stream (c->side), created on first use
int ensure_side(cqr_ctx* c);
int ensure_ws(cqr_ctx* c, size_t bytes);
template <class T>
inline T* ws_at(cqr_ctx* c, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(c->ws) + off);
}

// ---------------------------------------------------------------------------
// TSQR plan: leaves are contiguous row blocks; the tree pairs them with the
// reference's schedule (alg/QR/tsqr/bitree_tsqr.cxx:172-208).
// ---------------------------------------------------------------------------
struct TsqrPlan {
  int64_t m = 0, b = 0;
  int64_t nleaves = 0;
  int64_t max_leaf_rows = 0;
  std::vector<int64_t> leaf_off;     // nleaves + 1 row offsets
  // levels in forward order; each level is a list of (receiver, sender)
  std::vector<std::vector<std::pair<int, int>>> levels;
  int64_t nnodes = 0;                // total tree nodes
  std::vector<int64_t> level_node0;  // first node index of each level
};
int64_t leaf_capacity_rows(int64_t b);          // largest leaf one CU holds
int make_plan(int64_t m, int64_t b, int64_t nleaves_hint, int64_t leaf_rows,
              TsqrPlan& p);

// ---------------------------------------------------------------------------
// Batched small GEMM addressing: batch index z selects each operand as
//   BM_STRIDED  p + z·zs          (per-leaf / per-node b×b arrays)
//   BM_LEAFROWS p + rowoff(z)     (leaf z's rows inside a column-major panel)
//   BM_SHIFTED  p + (z+zoff)·zs   (the sender of tree pair z)
//   BM_LEAFROWS_SHIFTED p + rowoff(z+zoff)  (the sender's rows of tree pair z)
// m < 0 or k < 0 stands for the rows of leaf z.
// ---------------------------------------------------------------------------
//   BM_PROB_Y   p + bp[z].yoff + prow, ld bp[z].m   (problem z's Q1 block)
//   BM_PROB_A   p + bp[z].aoff + prow, ld bp[z].m   (problem z's A block)
//   BM_PROB_T   p + bp[z].toff                     (problem z's Tᵀ block)
enum { BM_STRIDED = 0, BM_LEAFROWS = 1, BM_SHIFTED = 2, BM_PROB_Y = 3, BM_PROB_A = 4, BM_PROB_T = 5,
       BM_LEAFROWS_SHIFTED = 6 };
struct BMat {
  const double* p;
  int64_t ld;
  int64_t zs;
  int mode;
};
struct BgemmArgs {
  BMat A, B, C;
  int m, n, k;            // m < 0 / k < 0: rows of leaf z
  int ta, tb, smA, smB, smC;
  double alpha, beta;
  int64_t pm;             // panel rows for BM_LEAFROWS / m < 0
  int pL;                 // leaves
  int zoff;               // for BM_SHIFTED
};
inline BMat bmat(const double* p, int64_t ld, int64_t zs = 0, int mode = BM_STRIDED) {
  BMat b; b.p = p; b.ld = ld; b.zs = zs; b.mode = mode; return b;
}

// ---------------------------------------------------------------------------
// Batched TSQR + Householder reconstruction: one row per problem of a group
// that shares the panel width b.  All offsets index the group's packed
// buffers: the problem's A block starts at Abase + aoff (column-major, ld
// lda = m), its leaves are global leaves leaf0.., its tree nodes global
// nodes node0.., its construct_Q1 items item0.., its Q1 block starts at
// row offset yoff (doubles) of the Y scratch (ld m), its Tᵀ at toff.
// The leaf kernels find their problem by binary search on leaf0; the tree
// and construct_Q1 launches map their blocks through a BatchSeg table.
// ---------------------------------------------------------------------------
struct BatchProb {
  int64_t aoff, m, yoff, toff;
  int L, leaf0, node0, item0;
};
enum { BP_LEAF = 0, BP_NODE = 1, BP_ITEM = 2 };
__host__ __device__ inline int bp_first(const BatchProb& P, int which) {
  return which == BP_LEAF ? P.leaf0 : which == BP_NODE ? P.node0 : P.item0;
}
// problem owning global index idx (bp_first non-decreasing in p)
__device__ inline int bp_find(const BatchProb* bp, int nprob, int idx, int which) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bp_first(bp[mid], which) <= idx) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Dispatch order of a batched launch with dependencies between its blocks
// (the tree nodes, construct_Q1's items): global block ids g0.. of a
// segment are the problem's local nodes (items) local0.., one level of one
// problem per segment.  Segments run level by level across all problems
// (children before parents for the tree, the root first for construct_Q1),
// so a block waits only on blocks of earlier levels, which every problem
// dispatched before it, instead of on a deep chain inside one problem while
// the other problems' ready work sits undispatched.
struct BatchSeg {
  int g0, prob, local0, n;
};
// segment holding global block idx (g0 increasing)
__device__ inline int seg_find(const BatchSeg* s, int nseg, int idx) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s[mid].g0 <= idx) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

#define CQR_MAX_TREE_NODES 8192
#define CQR_TRSM_MAX_BLOCKS 4096   // 128-row blocks of the one-launch triangular solve

// Column-pipelined tree over all levels (ptree_kernel.hip): node outputs R
// in nodeR (nnodes×b×b), V2/T/tau per node in forward level order.
int leaf_tree(cqr_ctx* c, double* A, int64_t lda, int64_t m, int L, int b, double* leafR, double* leafT,
              double* leafTau, double* nodeR, double* V2, double* T, double* tau, int nnodes);
int launch_ptree(cqr_ctx* c, const double* leafR, double* nodeR, double* V2, double* T, double* tau,
                 int L, int nnodes, int b);

// construct_Q1 over the whole tree and the leaves in one launch
// (q1_kernel.hip): C is scratch of L×b×b; Q receives the m×b Q1 (ld ldq).
int launch_q1(cqr_ctx* c, const double* nodeV2, const double* nodeT, const double* leafT, const double* V,
              int64_t ldv, int64_t m, double* C, double* Q, int64_t ldq, int L, int nnodes, int b);

// Batched launchers (one group of problems sharing b)
int leaf_qr_batched(cqr_ctx* c, double* Abase, int64_t total_m, int L, int64_t rows, int b, double* R, double* T,
                    double* tau, const BatchProb* bp, int nprob);
int launch_ptree_batched(cqr_ctx* c, const double* leafR, double* nodeR, double* V2, double* T, double* tau,
                         uint64_t* cnt, int total_nodes, int b, const BatchProb* bp, const BatchSeg* seg,
                         int nseg);
int launch_q1_batched(cqr_ctx* c, const double* nodeV2, const double* nodeT, const double* leafT, const double* Vbase,
                      double* C, double* Y, uint64_t* flags, int total_items, int b, const BatchProb* bp,
                      const BatchSeg* seg, int nseg);
int q1_items_per_problem(int L, int b);
// Batched Y(b:m) = Q1(b:m)·U⁻¹ of every problem (gemm_update.hip): one
// block per batched_y_tile_rows() rows of a problem, mapped by `seg`.
int batched_y_tile_rows();
int update_y_batched(cqr_ctx* c, const BatchProb* bp, const BatchSeg* seg, int nseg, int total_tiles, int b,
                     const double* Y, const double* Uinv, double* A);
int hr_core_batched(cqr_ctx* c, const BatchProb* bp, int nprob, const double* Y, const double* leafR,
                    const double* nodeR, double* Abase, double* W, double* Uinv, double* Winv, double* LinvT, int b);
int gemm_f64_batched_prob(cqr_ctx* c, int ta, int tb, int m, int n, int k, double alpha, const BMat& A, int smA,
                          const BMat& B, int smB, double beta, const BMat& C, int smC, int nprob, const BatchProb* bp,
                          int64_t prow, int m_prob, int max_m);

// Kernel launchers (tsqr_kernels.hip)
int leaf_qr(cqr_ctx* c, double* A, int64_t lda, int64_t m, int L, int b, double* R, double* T,
            double* tau);
int tree_level(cqr_ctx* c, double* Rleaf, int n, int node0, int b, double* V2, double* T,
               double* tau);
int bgemm(cqr_ctx* c, const BgemmArgs& g, int batch, int max_m);
// With Tt (b = 128 only): also Tᵀ, and U⁻¹·Tᵀ into M2 when non-null, and the
// top b rows of the explicit Y and of Y·Tᵀ into Ye / YT (ld ldy) when non-null.
int hr_core(cqr_ctx* c, const double* Q1top, int64_t ldq, const double* Rroot, double* Atop,
            int64_t lda, double* W, double* Uinv, double* Winv, double* LinvT, int64_t* signs, int b,
            double* Tt = nullptr, double* M2 = nullptr, double* Ye = nullptr, double* YT = nullptr,
            int64_t ldy = 0);
int launch_diag(cqr_ctx* c, const double* T, int64_t ldt, double* tau, int b);
int launch_signed_nlu(cqr_ctx* c, double* A, int64_t lda, double* R, int64_t lda_R,
                      int64_t* signs, int b);
int launch_tri_inverse(cqr_ctx* c, const double* X, int64_t ldx, double* Xinv,
                       int64_t ldi, int b, int upper, int unit);
// every blk-wide diagonal block of X's upper triangle (n×n) in one launch;
// block p's inverse at Xinv + p·blk² with ld bj (blk ≤ 128)
int launch_tri_inverse_diag(cqr_ctx* c, const double* X, int64_t ldx, double* Xinv, int64_t n, int blk);
int launch_tau_recon(cqr_ctx* c, char s, int64_t ncol, int64_t nrow, int64_t lda,
                     const double* Y, double* tau);
int launch_pack(cqr_ctx* c, int op, const double* src, double* dst, int64_t b,
                int64_t lda, int has_diag);
int launch_copy_unit_lower(cqr_ctx* c, const double* YR, int64_t ldyr, double* Y,
                           int64_t ldy, int64_t m, int64_t b);
int launch_blocked_T(cqr_ctx* c, const double* Tfull, double* Tblk, int b, int nb);
int launch_finite_scan(cqr_ctx* c, const double* A, int64_t lda, int64_t m, int64_t n);

// Structural masks for GEMM operands (stored-matrix coordinates)
enum { SM_NONE = 0, SM_UNIT_LOWER = 1, SM_UPPER = 2, SM_LOWER = 3, SM_IDENTITY = 4, SM_UNIT_UPPER = 5 };

// FP64 MFMA GEMMs (gemm_f64.hip)
//   C := alpha·op(A)·op(B) + beta·C, column-major; A's stored top-left square
//   may carry a structural mask (e.g. the implicit unit-lower Y of a panel).
int gemm_f64(cqr_ctx* c, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
             const double* A, int64_t lda, int smA, const double* B, int64_t ldb,
             double beta, double* C, int64_t ldc);
// Same with an explicit split-K scratch (nullptr: no split) and a mask on B.
int gemm_f64_ws(cqr_ctx* c, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
                const double* A, int64_t lda, int smA, const double* B, int64_t ldb, int smB,
                double beta, double* C, int64_t ldc, double* scratch);
// Trailing update of one panel: A (m×n) −= Y·Tᵀ·(Yᵀ·A), Y unit lower in its
// top b×b (stored as [Y\R] in the panel), Tt = Tᵀ (b×b, ld b).
int wy_update_f64(cqr_ctx* c, const double* Y, int64_t ldy, int y_implicit,
                  const double* Tt, double* A, int64_t lda, int64_t m, int64_t n, int b);
// Same with workspace (wy_update_ws_doubles(m, n, b) doubles), the choice of
// Qᵀ (t_trans = 0: Z' = Tᵀ·Z) or Q (t_trans = 1: Z' = T·Z), and the
// arithmetic of the two large products (CQR_PREC_*).
size_t wy_update_ws_doubles(int64_t m, int64_t n, int b);
int wy_update_ws(cqr_ctx* c, const double* Y, int64_t ldy, int y_implicit, const double* Tt, int t_trans,
                 double* A, int64_t lda, int64_t m, int64_t n, int b, double* ws, int prec);
int wy_update_f64_ws(cqr_ctx* c, const double* Y, int64_t ldy, int y_implicit, const double* Tt,
                     double* A, int64_t lda, int64_t m, int64_t n, int b, double* ws);
size_t gemm_f64_scratch_doubles(int64_t m, int64_t n, int64_t k);
// C (m×n) := alpha·op(A)·op(B) + beta·C for n ≤ 16 and m, k ≤ 256 in one
// workgroup; C may alias B (B is read completely first).
int launch_small_thin(cqr_ctx* c, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                      int64_t lda, int smA, const double* B, int64_t ldb, int smB, double beta, double* C,
                      int64_t ldc, int kind);
// A (m×n) −= YT·(Yeᵀ·A) with an explicit unit-lower Ye and YT = Ye·Tt (both
// m×b, ld m), the two products in `prec`; workspace wy_update_ws_doubles(m,
// n, b) doubles.
int wy_update_explicit(cqr_ctx* c, const double* Ye, const double* YT, int64_t m, int64_t n, int b, double* A,
                       int64_t lda, double* ws, int prec);
// Its two halves: Z = Yeᵀ·A into ws (*Zout points at it), then A −= YT·Z.
int wy_explicit_yta(cqr_ctx* c, const double* Ye, int64_t m, int64_t n, int b, const double* A, int64_t lda,
                    double* ws, int prec, double** Zout);
int wy_explicit_apply(cqr_ctx* c, const double* YT, int64_t m, int64_t n, int b, const double* Z, double* A,
                      int64_t lda, int prec);

// Least-squares kernels (lsq_kernels.hip): C = alpha·op(A)·X + beta·C for
// X with nr ≤ 16 columns; the transposed form needs
// thin_gemm_scratch_doubles(m, k, nr) of scratch and, with C = nullptr,
// only leaves its slab partials there (for launch_slab_reduce_t).  The
// plain form given scratch (and unit_top = 0) splits the columns of a large
// product over thin_n_splits(m, k) workgroup rows; the scratch then holds
// thin_n_scratch_doubles(m, k, nr).
int64_t thin_t_slab_rows(int64_t m, int64_t k);
int thin_n_splits(int64_t m, int64_t k);                      // column chunks of a plain A·X
size_t thin_n_scratch_doubles(int64_t m, int64_t k, int nr);  // its partial sums
size_t thin_gemm_scratch_doubles(int64_t m, int64_t k, int nr);
// unit_top > 0: A's first unit_top rows are unit lower (A(r, c) = 0 for
// c > r and 1 at c = r, whatever is stored there).
int thin_gemm(cqr_ctx* c, int ta, int64_t m, int64_t k, double alpha, const double* A, int64_t lda,
              const double* X, int64_t ldx, int nr, double beta, double* C, int64_t ldc, double* scratch,
              int unit_top = 0);
// Z' = op(Tt)·Σ slab partials of thin_gemm(ta = 1, C = nullptr) (k ≤ 128,
// nr ≤ 4), one workgroup
int launch_slab_reduce_t(cqr_ctx* c, const double* P, int nslab, int k, int nr, const double* Tt, int64_t ldt,
                         int ta, double* Zp, int64_t ldz);
// B(0:n, :) := R⁻¹·B or R⁻ᵀ·B (nr ≤ 16 columns) in one launch, with the
// 128-wide diagonal-block inverses Rinv
int launch_trsm_chain(cqr_ctx* c, const double* R, int64_t ldr, int64_t n, const double* Rinv, double* B,
                      int64_t ldb, int nr, int trans);
// One step of the blocked R solve (lsq_kernels.hip): blocks [q0, q0+nq)
// take block p's solution into account; block `solve` is then finished.
int launch_trsm_step(cqr_ctx* c, const double* R, int64_t ldr, int64_t n, const double* Rinv, double* B, int64_t ldb,
                     int nr, int trans, int p, int q0, int nq, int solve);
int launch_axpby(cqr_ctx* c, int64_t m, int64_t n, double alpha, const double* X, int64_t ldx, double beta,
                 double* Y, int64_t ldy);
int launch_colnorm(cqr_ctx* c, const double* X, int64_t ldx, int64_t m, int n, double* out);
// out[0..1] := max and min |R(i,i)| over i < n (NaN if any entry is NaN).
int launch_diag_range(cqr_ctx* c, const double* R, int64_t ldr, int64_t n, double* out);
int launch_copy2d(cqr_ctx* c, const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m,
                  int64_t n);
// dst (m×n, ld ldd) := v everywhere (a kernel, so graph capture records no memset)
int launch_fill2d(cqr_ctx* c, double* dst, int64_t ldd, int64_t m, int64_t n, double v);

// Drivers (qr_driver.cpp)
size_t panel_ws_bytes(int64_t m, int64_t b, int64_t nleaves, int64_t leaf_rows);
int hh_recon_panel(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t b, double* W, double* T,
                   double* invT_T, double* Tt, int64_t nleaves, char* base, int64_t* signs_out,
                   double* Ye = nullptr, double* YT = nullptr);
#define CQR_CHECK_CTX(c)                     \
  do {                                       \
    if (!(c)) return CQR_ERR_INVALID;        \
    (void)hipSetDevice((c)->device);         \
  } while (0)
int finite_check(cqr_ctx* c, const double* A, int64_t lda, int64_t m, int64_t n);
// QR_2D workspace layout (byte offsets into one arena) and the driver; the
// arena base is passed so callers can place their own buffers behind it.
struct Qr2dWS {
  size_t panel, upd_near, upd_far, W, Tt, Ye;
};
size_t qr2d_ws(int64_t m, int64_t k, int64_t b, int64_t leaf_rows, Qr2dWS* o);
int qr2d_run(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t k, int64_t b, double* aggreg_Y,
             int64_t lda_aY, char* base, const Qr2dWS& o, double* Tt_out, int prec);
struct LsqWS {
  size_t Af, Tt, Ri, R, F, G, H, Tmp, Q, N;
};
size_t lstsq_layout(int64_t m, int64_t n, int64_t nr, int64_t b, int64_t leaf_rows, Qr2dWS* o, LsqWS* w);
// The refinement's buffers alone (Af and Tt given by the caller).
size_t refine_ws(int64_t m, int64_t n, int64_t nr, int64_t b, LsqWS* w);
struct LsqRun {
  int iters = 0;
  double rel = 0.0;
};
// The solve and refinement of cqr_lstsq from given factors: [Y\R] of A in
// Af (ld ldf) and the Tᵀ blocks of its width-b panels in Tt; the buffers of
// w (refine_ws) at base.  Returns CQR_OK when the last relative update is
// ≤ tol (or max_iter is 0), CQR_ERR_NOCONVERGE when the refinement ended
// above tol, CQR_ERR_NONFINITE when the iterate or R holds a NaN or Inf,
// and −1 (only with `escalate`, for a mixed `prec`) when R's diagonal range
// already rules the mixed factors out.
int lstsq_refine(cqr_ctx* c, const double* A, int64_t lda, int64_t m, int64_t n, const double* B, int64_t ldb,
                 int64_t nr, double* X, int64_t ldx, int64_t b, const double* Af, int64_t ldf, const double* Tt,
                 int prec, int max_iter, double tol, bool escalate, char* base, const LsqWS& w, LsqRun& out);
int hh_recon_core(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t b, double* W,
                  double* T, double* invT_T, double* Tt, int64_t nleaves);
// Looks up the device address of the fixed scratch array of noalloc.hip
// (device current); cqr_create calls it once per context.
int noalloc_setup(cqr_ctx* c);
// dgeqrf of one m×b block on caller scratch of local_tsqr_ws_bytes bytes.
size_t local_tsqr_ws_bytes(int64_t m, int64_t b, int64_t leaf_rows);
int local_tsqr_ws(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t b, double* tau, double* T, char* base);
// bitree_tsqr without tree_data over P ranks (tsqr_apply.hip): each rank's
// row block gets its dgeqrf [V\R] in place, the tree reduces the ranks' R
// with the reference's schedule, and with output_Y the senders' top
// triangles receive the node reflectors V2 and the receivers' the node R.
int bitree_ranks(cqr_ctx* c, double* A, int64_t lda, double* R, double* tau, int64_t m, int64_t b, int64_t P,
                 int output_Y);
// Ranks doing work for P ranks over m rows of width b (bitree_tsqr.cxx:123-126).
inline int64_t ranks_working(int64_t m, int64_t b, int64_t P) { return P * b > m ? m / b : P; }

}  // namespace cqr
