// Drop-in C++ entry points with the reference's signatures
// (include/candmc/CANDMC_QR.h), served by the C-ABI of candmc_qr.h.
// Host pointers are staged through device buffers; device pointers are used
// in place.  Errors abort with a message, as the reference's asserts do.
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "../../include/candmc/CANDMC_QR.h"
#include "cqr_internal.h"

namespace {

std::mutex g_mu;
cqr_ctx* g_ctx = nullptr;

[[noreturn]] void die(const char* fn, const char* msg) {
  fprintf(stderr, "candmc_qr: %s: %s\n", fn, msg);
  abort();
}

void check(int s, const char* fn) {
  if (s != CQR_OK) {
    char buf[600];
    snprintf(buf, sizeof buf, "%s (%s)", cqr_status_string(s), g_ctx ? cqr_last_error(g_ctx) : "");
    die(fn, buf);
  }
}

void need_single_rank(const CommData_t& c, const char* fn) {
  if (c.np > 1) die(fn, "multi-rank communicators are not served on one GPU (np must be <= 1)");
}

bool on_device(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Device view of a caller buffer of `n` doubles.
struct Stage {
  double* host = nullptr;
  double* dev = nullptr;
  size_t n = 0;
  bool owned = false;
  bool out = false;
  Stage(const double* p, size_t count, bool is_out, const char* fn) : n(count), out(is_out) {
    if (!p) return;
    if (on_device(p)) {
      dev = const_cast<double*>(p);
      return;
    }
    host = const_cast<double*>(p);
    if (hipMalloc(&dev, n * sizeof(double)) != hipSuccess) die(fn, "cannot stage a host buffer");
    owned = true;
    if (hipMemcpy(dev, host, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      die(fn, "host-to-device copy failed");
  }
  void finish(const char* fn) {
    if (owned && out && hipMemcpy(host, dev, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      die(fn, "device-to-host copy failed");
  }
  ~Stage() {
    if (owned) (void)hipFree(dev);
  }
};

struct Stage64 {
  int64_t* host = nullptr;
  int64_t* dev = nullptr;
  size_t n;
  bool owned = false;
  Stage64(int64_t* p, size_t count, const char* fn) : n(count) {
    if (!p) return;
    if (on_device(p)) {
      dev = p;
      return;
    }
    host = p;
    if (hipMalloc(&dev, n * sizeof(int64_t)) != hipSuccess) die(fn, "cannot stage signs");
    owned = true;
  }
  void finish(const char* fn) {
    if (owned && hipMemcpy(host, dev, n * sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess)
      die(fn, "device-to-host copy failed");
  }
  ~Stage64() {
    if (owned) (void)hipFree(dev);
  }
};

void sync(const char* fn) { check(cqr_sync(g_ctx), fn); }

}  // namespace

cqr_ctx* candmc_qr_context() {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ctx) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (cqr_create(&g_ctx, dev) != CQR_OK) die("candmc_qr_context", "no usable GPU");
  }
  return g_ctx;
}

// ---------------------------------------------------------------- tsqr
void tree_tsqr(double* A, double* buf, int64_t const b, double* tau) {
  (void)buf;
  const char* fn = "tree_tsqr";
  cqr_ctx* c = candmc_qr_context();
  Stage a(A, (size_t)2 * b * b, true, fn), t(tau, (size_t)std::min<int64_t>(b, TAU_BLK) * b, true, fn);
  check(cqr_tree_tsqr(c, a.dev, b, t.dev), fn);
  sync(fn);
  a.finish(fn);
  t.finish(fn);
}

void local_tsqr(double* A, double* R, int64_t const m, int64_t const b, int64_t const lda_A, double* tau) {
  (void)R;
  const char* fn = "local_tsqr";
  cqr_ctx* c = candmc_qr_context();
  Stage a(A, (size_t)lda_A * b, true, fn);
  std::vector<double> tmp;
  if (!tau) {
    tmp.resize(b);
    tau = tmp.data();
  }
  Stage t(tau, (size_t)b, true, fn);
  check(cqr_local_tsqr(c, a.dev, lda_A, m, b, t.dev, nullptr), fn);
  sync(fn);
  a.finish(fn);
  t.finish(fn);
}

namespace {

// ---- tree factors of the several-rank path ------------------------------
// The library's TSQR tree over numPes virtual ranks keeps per-leaf and
// per-node factors (cqr_tree_data_doubles doubles) and every leaf's taus,
// more than the reference sizes its tree_data (2b²(ln P + 2) doubles,
// alg/QR/hh_recon/hh_recon.cxx:151) and tau (b doubles) for.  So the
// factors stay on the device, held here, and the caller's tree_data
// receives only a five-double tag {magic, id, m, b, numPes} that names
// them; construct_Q1 and apply_tsqr_QT find them by that tag.  The last
// kTreeCache factorizations are held (the oldest is dropped first).
constexpr double kTreeMagic = 7.0e300;
constexpr size_t kTreeCache = 8;
constexpr int kTagDoubles = 5;
struct TreeEntry {
  int64_t id, m, b, P;
  double* td = nullptr;   // cqr_tree_data_doubles(m, b, P) doubles
  double* tau = nullptr;  // every leaf's taus
};
std::vector<TreeEntry> g_trees;
int64_t g_tree_next = 1;

void put_tag(double* dst, const double* tag, const char* fn) {
  if (on_device(dst)) {
    if (hipMemcpy(dst, tag, kTagDoubles * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      die(fn, "cannot write the tree_data tag");
  } else {
    for (int i = 0; i < kTagDoubles; ++i) dst[i] = tag[i];
  }
}

const TreeEntry& find_tree(const double* tree_data, int64_t m, int64_t b, int64_t P, const char* fn) {
  double tag[kTagDoubles];
  if (on_device(tree_data)) {
    if (hipMemcpy(tag, tree_data, sizeof tag, hipMemcpyDeviceToHost) != hipSuccess) die(fn, "cannot read tree_data");
  } else {
    for (int i = 0; i < kTagDoubles; ++i) tag[i] = tree_data[i];
  }
  if (tag[0] != kTreeMagic) die(fn, "tree_data does not come from bitree_tsqr in this process");
  std::lock_guard<std::mutex> lk(g_mu);
  for (const TreeEntry& e : g_trees)
    if ((double)e.id == tag[1]) {
      if (e.m != m || e.b != b || e.P != P) die(fn, "tree_data belongs to a TSQR of another shape");
      return e;
    }
  die(fn, "the TSQR factors of this tree_data were released (more than 8 newer factorizations)");
}

}  // namespace

// numPes virtual ranks.  tree_data == NULL, or a single working rank: the
// reference's rank semantics (cqr_bitree_tsqr without tree_data: dgeqrf per
// rank, V sent back with output_Y, tree_data untouched).  Otherwise the
// library's leaves: A := leaf [V\R] blocks, the factors held on the device
// and tree_data := their tag.  tau always receives b doubles (rank 0's, or
// the first leaf's).
void bitree_tsqr(double* A, int64_t const lda_A, double* R, double* tau, int64_t const m, int64_t const b,
                 int64_t const myRank, int64_t const numPes, int64_t const root, int64_t const req_id,
                 CommData_t cdt, int64_t const output_Y, double* tree_data) {
  (void)myRank; (void)root; (void)req_id;
  const char* fn = "bitree_tsqr";
  need_single_rank(cdt, fn);
  if (numPes < 1 || b < 1 || m < b || lda_A < m) die(fn, "bad arguments (needs numPes >= 1, m >= b, lda_A >= m)");
  cqr_ctx* c = candmc_qr_context();
  std::vector<double> tmp;
  if (!tau) {  // the reference's local_tsqr then uses a buffer of its own
    tmp.resize(b);
    tau = tmp.data();
  }
  Stage a(A, (size_t)lda_A * b, true, fn), r(R, (size_t)b * b, true, fn), t(tau, (size_t)b, true, fn);
  if (!tree_data || cqr::ranks_working(m, b, numPes) == 1) {
    check(cqr_bitree_tsqr(c, a.dev, lda_A, r.dev, t.dev, m, b, numPes, (int)output_Y, nullptr), fn);
    sync(fn);
    a.finish(fn);
    r.finish(fn);
    t.finish(fn);
    return;
  }
  int64_t L = 0;
  check(cqr_plan_tsqr(m, b, numPes, 0, &L, nullptr, nullptr), fn);
  TreeEntry e;
  e.m = m; e.b = b; e.P = numPes;
  const size_t ntd = (size_t)cqr_tree_data_doubles(m, b, numPes);
  if (hipMalloc(&e.td, ntd * sizeof(double)) != hipSuccess || hipMalloc(&e.tau, (size_t)L * b * sizeof(double)) != hipSuccess)
    die(fn, "cannot hold the tree factors on the device");
  check(cqr_bitree_tsqr(c, a.dev, lda_A, r.dev, e.tau, m, b, numPes, (int)output_Y, e.td), fn);
  if (hipMemcpyAsync(t.dev, e.tau, (size_t)b * sizeof(double), hipMemcpyDeviceToDevice,
                     (hipStream_t)cqr_get_stream(c)) != hipSuccess)
    die(fn, "tau copy failed");
  sync(fn);
  a.finish(fn);
  r.finish(fn);
  t.finish(fn);
  double tag[kTagDoubles];
  {
    std::lock_guard<std::mutex> lk(g_mu);
    e.id = g_tree_next++;
    if (g_trees.size() >= kTreeCache) {
      (void)hipFree(g_trees.front().td);
      (void)hipFree(g_trees.front().tau);
      g_trees.erase(g_trees.begin());
    }
    g_trees.push_back(e);
  }
  tag[0] = kTreeMagic; tag[1] = (double)e.id; tag[2] = (double)m; tag[3] = (double)b; tag[4] = (double)numPes;
  put_tag(tree_data, tag, fn);
}

void construct_Q1(double const* Y, int64_t const lda_Y, double const* tau, double* Q1, int64_t const lda_Q,
                  int64_t const m, int64_t const b, int64_t const k, int64_t const myRank, int64_t const numPes,
                  int64_t const root, CommData_t cdt, double* tree_data, double* ID) {
  (void)myRank; (void)root;
  const char* fn = "construct_Q1";
  need_single_rank(cdt, fn);
  if (numPes < 1 || b < 1 || m < b || k < 1 || k > m) die(fn, "bad arguments");
  cqr_ctx* c = candmc_qr_context();
  const bool one = cqr::ranks_working(m, b, numPes) == 1;
  if (!one && !tree_data) die(fn, "tree_data from bitree_tsqr is required for several ranks");
  Stage y(Y, (size_t)lda_Y * b, false, fn), q(Q1, (size_t)lda_Q * k, true, fn), id(ID, (size_t)b * k, false, fn);
  // ID is b×k with leading dimension b (construct_tsqr_Q.cxx:99,125)
  if (one) {  // the one-rank dorgqr / dormqr branch (construct_tsqr_Q.cxx:87-102)
    Stage t(tau, (size_t)b, false, fn);
    check(cqr_construct_Q1_id(c, y.dev, lda_Y, t.dev, q.dev, lda_Q, m, b, k, 1, nullptr, id.dev, b), fn);
  } else {
    const TreeEntry& e = find_tree(tree_data, m, b, numPes, fn);
    check(cqr_construct_Q1_id(c, y.dev, lda_Y, e.tau, q.dev, lda_Q, m, b, k, numPes, e.td, id.dev, b), fn);
  }
  sync(fn);
  q.finish(fn);
}

void apply_tsqr_QT(double const* Y, int64_t const lda_Y, double const* tau, double* B, int64_t const lda_B,
                   int64_t const m, int64_t const b, int64_t const k, int64_t const myRank, int64_t const numPes,
                   int64_t const root, CommData_t cdt, double* tree_data) {
  (void)myRank; (void)root;
  const char* fn = "apply_tsqr_QT";
  need_single_rank(cdt, fn);
  if (numPes < 1 || b < 1 || m < b || k < 0) die(fn, "bad arguments");
  cqr_ctx* c = candmc_qr_context();
  const bool one = cqr::ranks_working(m, b, numPes) == 1;
  if (!one && !tree_data) die(fn, "tree_data from bitree_tsqr is required for several ranks");
  Stage y(Y, (size_t)lda_Y * b, false, fn), bm(B, (size_t)lda_B * k, true, fn);
  if (one) {  // dormqr('L','T') of the one-rank branch (apply_tsqr_QT.cxx:73-80)
    Stage t(tau, (size_t)b, false, fn);
    check(cqr_apply_tsqr_QT(c, y.dev, lda_Y, t.dev, bm.dev, lda_B, m, b, k, 1, nullptr), fn);
  } else {
    const TreeEntry& e = find_tree(tree_data, m, b, numPes, fn);
    check(cqr_apply_tsqr_QT(c, y.dev, lda_Y, e.tau, bm.dev, lda_B, m, b, k, numPes, e.td), fn);
  }
  sync(fn);
  bm.finish(fn);
}

// One rank per communicator: the one-rank branches of bitree_tsqr and
// apply_tsqr_QT (dgeqrf of each panel, dormqr of the trailing columns).
void QR_tree_2D(double* A, int64_t const lda_A, int64_t const m, int64_t const k, int64_t const b,
                int64_t const myRank, int64_t const numPes, int64_t const root_row, int64_t const root_col,
                CommData_t cdt_row, CommData_t cdt_col, CommData_t cdt_world, int64_t const _stop_at) {
  (void)myRank; (void)numPes; (void)root_row; (void)root_col;
  const char* fn = "QR_tree_2D";
  need_single_rank(cdt_row, fn);
  need_single_rank(cdt_col, fn);
  need_single_rank(cdt_world, fn);
  cqr_ctx* c = candmc_qr_context();
  Stage a(A, (size_t)lda_A * k, true, fn);
  check(cqr_qr_tree_2d(c, a.dev, lda_A, m, k, b, _stop_at, 1), fn);
  sync(fn);
  a.finish(fn);
}

// ---------------------------------------------------------------- hh_recon
void compute_invT_from_W(double const* W, int64_t b, double* invT) {
  const char* fn = "compute_invT_from_W";
  cqr_ctx* c = candmc_qr_context();
  Stage w(W, (size_t)b * b, false, fn), t(invT, (size_t)b * b, true, fn);
  check(cqr_compute_invT_from_W(c, w.dev, b, t.dev), fn);
  sync(fn);
  t.finish(fn);
}

static void nlu(bool rec, double* A, double* R, int64_t b, int64_t lda_A, int64_t lda_R, int64_t* signs) {
  const char* fn = rec ? "recursive_NLU" : "signed_NLU";
  cqr_ctx* c = candmc_qr_context();
  Stage a(A, (size_t)lda_A * b, true, fn), r(R, (size_t)lda_R * b, true, fn);
  Stage64 s(signs, (size_t)b, fn);
  check((rec ? cqr_recursive_nlu : cqr_signed_nlu)(c, a.dev, lda_A, r.dev, lda_R, b, s.dev), fn);
  sync(fn);
  a.finish(fn);
  r.finish(fn);
  s.finish(fn);
}

void signed_NLU(double* A, double* R, int64_t b, int64_t lda_A, int64_t lda_R, int64_t* signs) {
  nlu(false, A, R, b, lda_A, lda_R, signs);
}

void recursive_NLU(double* A, double* R, int64_t b, int64_t lda_A, int64_t lda_R, int64_t* signs) {
  nlu(true, A, R, b, lda_A, lda_R, signs);
}

void hh_recon_qr(double* A, int64_t lda_A, int64_t m, int64_t b, double* W, int64_t myRank, int64_t numPes,
                 int64_t root, int64_t req_id, CommData_t cdt) {
  (void)myRank; (void)root; (void)req_id;
  const char* fn = "hh_recon_qr";
  need_single_rank(cdt, fn);
  cqr_ctx* c = candmc_qr_context();
  Stage a(A, (size_t)lda_A * b, true, fn), w(W, (size_t)b * b, true, fn);
  check(cqr_hh_recon_qr(c, a.dev, lda_A, m, b, w.dev, numPes), fn);
  sync(fn);
  a.finish(fn);
  w.finish(fn);
}

// ---------------------------------------------------------------- qr_2d
static void need_grid(const pview* pv, const char* fn) {
  if (pv && (pv->crow.np > 1 || pv->ccol.np > 1)) die(fn, "the process grid must be 1x1 on one GPU");
}

void compute_invT_from_Y(double const* Y, int64_t lda_Y, double* T, int64_t lda_T, int64_t mb, int64_t b,
                         pview* pv) {
  const char* fn = "compute_invT_from_Y";
  need_grid(pv, fn);
  cqr_ctx* c = candmc_qr_context();
  Stage y(Y, (size_t)lda_Y * b, false, fn), t(T, (size_t)lda_T * b, true, fn);
  check(cqr_compute_invT_from_Y(c, y.dev, lda_Y, t.dev, lda_T, mb, b), fn);
  sync(fn);
  t.finish(fn);
}

// The two-Y form (qr_2d.cxx:76-106): T (b1×b2) := Y1ᵀ·Y2.
void compute_invT_from_Y(double const* Y1, int64_t lda_Y1, double const* Y2, int64_t lda_Y2, double* T,
                         int64_t lda_T, int64_t mb, int64_t b1, int64_t b2, pview* pv) {
  const char* fn = "compute_invT_from_Y";
  need_grid(pv, fn);
  cqr_ctx* c = candmc_qr_context();
  Stage y1(Y1, (size_t)lda_Y1 * b1, false, fn), y2(Y2, (size_t)lda_Y2 * b2, false, fn),
      t(T, (size_t)lda_T * b2, true, fn);
  check(cqr_compute_invT_from_Y2(c, y1.dev, lda_Y1, y2.dev, lda_Y2, t.dev, lda_T, mb, b1, b2), fn);
  sync(fn);
  t.finish(fn);
}

void update_A(double const* Y, int64_t lda_Y, double* A, int64_t lda_A, int64_t m, int64_t k, int64_t b,
              double const* W, pview* pv, double* aggreg_Y, int64_t lda_aY, bool W_is_T) {
  const char* fn = "update_A";
  need_grid(pv, fn);
  cqr_ctx* c = candmc_qr_context();
  Stage y(Y, (size_t)lda_Y * b, false, fn), a(A, (size_t)lda_A * k, true, fn), w(W, (size_t)b * b, false, fn),
      ag(aggreg_Y, (size_t)lda_aY * b, true, fn);
  check(cqr_update_A(c, y.dev, lda_Y, a.dev, lda_A, m, k, b, w.dev, ag.dev, lda_aY, W_is_T ? 1 : 0), fn);
  sync(fn);
  a.finish(fn);
  ag.finish(fn);
}

void upd_A(double const* Ybuf, int64_t lda_Y, double* A, int64_t lda_A, int64_t mb, int64_t kb, int64_t b,
           double const* W, pview* pv, bool W_is_T) {
  const char* fn = "upd_A";
  need_grid(pv, fn);
  cqr_ctx* c = candmc_qr_context();
  Stage y(Ybuf, (size_t)lda_Y * b, false, fn), a(A, (size_t)lda_A * kb, true, fn), w(W, (size_t)b * b, false, fn);
  check(cqr_upd_A(c, y.dev, lda_Y, a.dev, lda_A, mb, kb, b, w.dev, W_is_T ? 1 : 0), fn);
  sync(fn);
  a.finish(fn);
}

void QR_2D(double* A, int64_t lda_A, int64_t m, int64_t k, int64_t b, pview* pv, double* aggreg_Y,
           int64_t lda_aY) {
  const char* fn = "QR_2D";
  need_grid(pv, fn);
  cqr_ctx* c = candmc_qr_context();
  Stage a(A, (size_t)lda_A * k, true, fn), ag(aggreg_Y, (size_t)lda_aY * k, true, fn);
  check(cqr_qr_2d(c, a.dev, lda_A, m, k, b, ag.dev, lda_aY), fn);
  sync(fn);
  a.finish(fn);
  ag.finish(fn);
  if (pv) {  // qr_2d.cxx:336-337 advances the roots once per panel
    pv->rrow = 0;
    pv->rcol = 0;
  }
}

void QR_2D_2D(double* A, int64_t lda_A, int64_t m, int64_t k, int64_t b, int64_t b_sub, pview* pv,
              double* aggreg_Y, int64_t lda_aY, int* desc_A, double* org_A, int64_t IA, int64_t JA) {
  (void)IA; (void)JA;
  const char* fn = "QR_2D_2D";
  need_grid(pv, fn);
  if (desc_A || org_A) die(fn, "the ScaLAPACK panel (desc_A) is not served");
  cqr_ctx* c = candmc_qr_context();
  Stage a(A, (size_t)lda_A * k, true, fn), ag(aggreg_Y, (size_t)lda_aY * k, true, fn);
  if (b >= m || b >= k || b == b_sub)  // the reference falls through to QR_2D (qr_2d.cxx:888-896)
    check(cqr_qr_2d(c, a.dev, lda_A, m, k, b_sub, ag.dev, lda_aY), fn);
  else
    check(cqr_qr_2d_2d(c, a.dev, lda_A, m, k, b, b_sub, ag.dev, lda_aY), fn);
  sync(fn);
  a.finish(fn);
  ag.finish(fn);
}

void QR_2D_pipe(double* A, int64_t lda_A, int64_t m, int64_t k, int64_t b, pview* pv, double* last_Y,
                int64_t lda_lY, double* last_W, double* my_last_W) {
  // The lookahead pipeline of qr_2d.cxx:378-635 produces the same [Y\R];
  // on one GPU the panel/update overlap is a scheduling matter.
  (void)last_Y; (void)lda_lY; (void)last_W; (void)my_last_W;
  QR_2D(A, lda_A, m, k, b, pv, nullptr, 0);
}

// ---------------------------------------------------------------- util
void tau_recon(const char structure, const int ncol, const int nrow, const int lda, double const* Y,
               double* tau) {
  const char* fn = "tau_recon";
  cqr_ctx* c = candmc_qr_context();
  Stage y(Y, (size_t)lda * ncol, false, fn), t(tau, (size_t)ncol, true, fn);
  check(cqr_tau_recon(c, structure, ncol, nrow, lda, y.dev, t.dev), fn);
  sync(fn);
  t.finish(fn);
}

void pack_upper(double const* R_in, double* R_out, int const b, int const lda_in) {
  const char* fn = "pack_upper";
  cqr_ctx* c = candmc_qr_context();
  Stage r(R_in, (size_t)lda_in * b, false, fn), p(R_out, (size_t)b * (b + 1) / 2, true, fn);
  check(cqr_pack_upper(c, r.dev, p.dev, b, lda_in), fn);
  sync(fn);
  p.finish(fn);
}

void unpack_upper(double const* R_in, double* R_out, int const b, int const lda_out) {
  const char* fn = "unpack_upper";
  cqr_ctx* c = candmc_qr_context();
  Stage p(R_in, (size_t)b * (b + 1) / 2, false, fn), r(R_out, (size_t)lda_out * b, true, fn);
  check(cqr_unpack_upper(c, p.dev, r.dev, b, lda_out), fn);
  sync(fn);
  r.finish(fn);
}

void pack_lower(double const* R_in, double* R_out, int const b, int const lda_in, int const has_diag) {
  const char* fn = "pack_lower";
  cqr_ctx* c = candmc_qr_context();
  Stage r(R_in, (size_t)lda_in * b, false, fn), p(R_out, (size_t)b * (b + 1) / 2, true, fn);
  check(cqr_pack_lower(c, r.dev, p.dev, b, lda_in, has_diag), fn);
  sync(fn);
  p.finish(fn);
}

void unpack_lower(double const* R_in, double* R_out, int const b, int const lda_out, int const has_diag) {
  const char* fn = "unpack_lower";
  cqr_ctx* c = candmc_qr_context();
  Stage p(R_in, (size_t)b * (b + 1) / 2, false, fn), r(R_out, (size_t)lda_out * b, true, fn);
  check(cqr_unpack_lower(c, p.dev, r.dev, b, lda_out, has_diag), fn);
  sync(fn);
  r.finish(fn);
}
