// Batched TSQR + Householder reconstruction of many independent tall-skinny
// problems (m_p × n_p, n_p ≤ 128, packed one after another in one buffer).
//
// Each problem is one panel of hh_recon_qr (alg/QR/hh_recon/hh_recon.cxx:
// 129-321) with b = n_p, so its result is [Y\R] in dgeqrf layout plus the
// compact-WY Tᵀ.  Problems sharing n form a group; a group runs as six
// launches over all of its problems (leaf QRs, the pipelined trees,
// construct_Q1, the HR cores and two batched GEMMs), each kernel finding its
// problem in a device table.  The whole sequence can be captured in a HIP
// graph and replayed.  The leaf height of every group is a tunable that
// cqr_batch_tune picks by timing the candidates on the caller's data.
#include <algorithm>
#include <cstring>
#include <vector>

#include "cqr_internal.h"

namespace cqr {

namespace {

struct Group {
  int b = 0;
  std::vector<int> probs;        // problem indices, in packing order
  int64_t leaf_rows = 0;         // 0 = automatic
  std::vector<BatchProb> hp;
  // block orders: tree (nseg_tree segments), construct_Q1 (nseg_q1), Y
  // product tiles (nseg_y, ytiles blocks)
  std::vector<BatchSeg> hseg;
  std::vector<char> htab;        // device table image: hp, then hseg
  int nseg_tree = 0, nseg_q1 = 0, nseg_y = 0, ytiles = 0;
  int total_leaves = 0, total_nodes = 0, total_items = 0;
  int64_t total_rows = 0, max_leaf_rows = 0, max_m = 0;
  BatchProb* d_bp = nullptr;     // the device table (d_seg follows it)
  BatchSeg* d_seg = nullptr;
  // byte offsets into the batch arena
  size_t leafR, leafT, leafTau, nodeR, nodeV2, nodeT, nodeTau, C, Y, W, Uinv, Winv, LinvT, cnt, qflags, bytes;
};

}  // namespace

int fill_u64(cqr_ctx* c, uint64_t* p, int64_t n);
int convert_f32_f64(cqr_ctx* c, const void* src, void* dst, int64_t n, int to_f64);

}  // namespace cqr

using namespace cqr;

struct cqr_batch {
  int64_t nprob = 0;
  std::vector<int64_t> m, n, aoff, toff;
  int64_t total_a = 0, total_t = 0;
  std::vector<Group> groups;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  double* stage = nullptr;       // FP64 copies of FP32 inputs (A then Tt)
  hipGraphExec_t gexec = nullptr;
  const void* g_A = nullptr;
  const void* g_T = nullptr;
  int g_dtype = -1;
  hipStream_t g_stream = nullptr;
  hipStream_t cap = nullptr;     // capture stream (the caller's may be the null stream)
};

namespace {

// Tree levels of an L-leaf problem: level k has pairs[k] nodes from node0[k].
void tree_levels(int L, std::vector<int>& node0, std::vector<int>& pairs) {
  node0.clear();
  pairs.clear();
  for (int n = L, first = 0; n > 1; n = n / 2 + n % 2) {
    node0.push_back(first);
    pairs.push_back(n / 2);
    first += n / 2;
  }
}

// Dispatch order of the group's tree nodes and construct_Q1 items: level by
// level across all problems (problems with the deepest trees first inside a
// level, as their chains are the longest), then construct_Q1's leaf items.
void build_order(Group& g) {
  const int np = (int)g.hp.size(), ncb = (g.b + 15) / 16;
  std::vector<int> order(np);
  for (int i = 0; i < np; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return g.hp[x].L > g.hp[y].L; });
  std::vector<std::vector<int>> n0(np), pr(np);
  int depth = 0;
  for (int i = 0; i < np; ++i) {
    tree_levels(g.hp[i].L, n0[i], pr[i]);
    depth = std::max(depth, (int)n0[i].size());
  }
  g.hseg.clear();
  int next = 0;
  for (int k = 0; k < depth; ++k)           // tree: children before parents
    for (int i : order)
      if (k < (int)n0[i].size()) {
        g.hseg.push_back(BatchSeg{next, i, n0[i][k], pr[i][k]});
        next += pr[i][k];
      }
  g.nseg_tree = (int)g.hseg.size();
  next = 0;
  for (int d = 0; d < depth; ++d)           // construct_Q1: the root level first
    for (int i : order) {
      const int D = (int)n0[i].size();
      if (d >= D) continue;
      const int k = D - 1 - d, nn = g.hp[i].L - 1;
      g.hseg.push_back(BatchSeg{next, i, (nn - n0[i][k] - pr[i][k]) * ncb, pr[i][k] * ncb});
      next += pr[i][k] * ncb;
    }
  for (int i : order) {                     // then every leaf
    const int L = g.hp[i].L;
    g.hseg.push_back(BatchSeg{next, i, (L - 1) * ncb, L * ncb});
    next += L * ncb;
  }
  g.nseg_q1 = (int)g.hseg.size() - g.nseg_tree;
  next = 0;
  const int64_t bm = batched_y_tile_rows();
  for (int i : order) {                     // Y product: the row tiles below the top b rows
    const int64_t rows = g.hp[i].m - g.b;
    if (rows <= 0) continue;
    const int t = (int)((rows + bm - 1) / bm);
    g.hseg.push_back(BatchSeg{next, i, 0, t});
    next += t;
  }
  g.nseg_y = (int)g.hseg.size() - g.nseg_tree - g.nseg_q1;
  g.ytiles = next;
  const size_t pb = sizeof(BatchProb) * g.hp.size(), sb = sizeof(BatchSeg) * g.hseg.size();
  g.htab.resize(pb + sb);
  std::memcpy(g.htab.data(), g.hp.data(), pb);
  std::memcpy(g.htab.data() + pb, g.hseg.data(), sb);
}

// Places the group's device table at d (g.htab.size() bytes).
void set_table(Group& g, void* d) {
  g.d_bp = static_cast<BatchProb*>(d);
  g.d_seg = reinterpret_cast<BatchSeg*>(static_cast<char*>(d) + sizeof(BatchProb) * g.hp.size());
}

// Plans a group's problems; with `alloc` its device table is allocated and
// filled here, without it the caller places (set_table) and fills it.
int plan_group(cqr_ctx* c, cqr_batch* bt, Group& g, bool alloc = true) {
  const int b = g.b;
  g.hp.clear();
  g.total_leaves = g.total_nodes = g.total_items = 0;
  g.total_rows = g.max_leaf_rows = g.max_m = 0;
  for (int p : g.probs) {
    TsqrPlan pl;
    const int s = make_plan(bt->m[p], b, 0, g.leaf_rows, pl);
    if (s != CQR_OK) return set_err(c, s, "batch: cannot plan problem %d (m=%lld, n=%d)", p, (long long)bt->m[p], b);
    BatchProb P{};
    P.aoff = bt->aoff[p];
    P.m = bt->m[p];
    P.yoff = g.total_rows * b;
    P.toff = bt->toff[p];
    P.L = (int)pl.nleaves;
    P.leaf0 = g.total_leaves;
    P.node0 = g.total_nodes;
    P.item0 = g.total_items;
    g.hp.push_back(P);
    g.total_leaves += P.L;
    g.total_nodes += P.L - 1;
    g.total_items += q1_items_per_problem(P.L, b);
    g.total_rows += P.m;
    g.max_leaf_rows = std::max(g.max_leaf_rows, pl.max_leaf_rows);
    g.max_m = std::max(g.max_m, P.m);
  }
  const size_t bb = (size_t)b * b, np = g.probs.size(), nn = std::max(g.total_nodes, 1);
  Sizer s;
  g.leafR = s.add(8 * g.total_leaves * bb);
  g.leafT = s.add(8 * g.total_leaves * bb);
  g.leafTau = s.add(8 * g.total_leaves * (size_t)b);
  g.nodeR = s.add(8 * nn * bb);
  g.nodeV2 = s.add(8 * nn * bb);
  g.nodeT = s.add(8 * nn * bb);
  g.nodeTau = s.add(8 * nn * (size_t)b);
  g.C = s.add(8 * g.total_leaves * bb);
  g.Y = s.add(8 * (size_t)g.total_rows * b);
  g.W = s.add(8 * np * bb);
  g.Uinv = s.add(8 * np * bb);
  g.Winv = s.add(8 * np * bb);
  g.LinvT = s.add(8 * np * bb);
  g.cnt = s.add(8 * nn);
  g.qflags = s.add(8 * nn * (size_t)((b + 15) / 16));
  g.bytes = s.bytes;
  build_order(g);
  if (!alloc) return CQR_OK;
  if (g.d_bp) (void)hipFree(g.d_bp);
  g.d_bp = nullptr;
  g.d_seg = nullptr;
  void* d = nullptr;
  CQR_HIP(c, hipMalloc(&d, g.htab.size()));
  set_table(g, d);
  CQR_HIP(c, hipMemcpy(d, g.htab.data(), g.htab.size(), hipMemcpyHostToDevice));
  return CQR_OK;
}

int ensure_batch_ws(cqr_ctx* c, cqr_batch* bt) {
  size_t need = 0;
  for (const Group& g : bt->groups) need = std::max(need, g.bytes);
  if (need <= bt->ws_bytes) return CQR_OK;
  CQR_HIP(c, hipStreamSynchronize(c->stream));
  if (bt->ws) CQR_HIP(c, hipFree(bt->ws));
  bt->ws = nullptr;
  bt->ws_bytes = 0;
  if (hipMalloc(&bt->ws, need) != hipSuccess) {
    (void)hipGetLastError();
    return set_err(c, CQR_ERR_NOMEM, "batch workspace of %zu bytes", need);
  }
  bt->ws_bytes = need;
  return CQR_OK;
}

// The six launches of one group on the context's stream.
int run_group(cqr_ctx* c, cqr_batch* bt, const Group& g, double* A, double* Tt) {
  char* base = static_cast<char*>(bt->ws);
  auto at = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
  const int b = g.b, np = (int)g.probs.size();
  const int64_t bb = (int64_t)b * b;
  uint64_t* cnt = reinterpret_cast<uint64_t*>(base + g.cnt);
  uint64_t* qf = reinterpret_cast<uint64_t*>(base + g.qflags);
  CQR_TRY(fill_u64(c, cnt, std::max(g.total_nodes, 1)));
  CQR_TRY(fill_u64(c, qf, (int64_t)std::max(g.total_nodes, 1) * ((b + 15) / 16)));
  CQR_TRY(leaf_qr_batched(c, A, g.total_rows, g.total_leaves, g.max_leaf_rows, b, at(g.leafR), at(g.leafT),
                          at(g.leafTau), g.d_bp, np));
  CQR_TRY(launch_ptree_batched(c, at(g.leafR), at(g.nodeR), at(g.nodeV2), at(g.nodeT), at(g.nodeTau), cnt,
                               g.total_nodes, b, g.d_bp, g.d_seg, g.nseg_tree));
  CQR_TRY(launch_q1_batched(c, at(g.nodeV2), at(g.nodeT), at(g.leafT), A, at(g.C), at(g.Y), qf, g.total_items, b,
                            g.d_bp, g.d_seg + g.nseg_tree, g.nseg_q1));
  CQR_TRY(hr_core_batched(c, g.d_bp, np, at(g.Y), at(g.leafR), at(g.nodeR), A, at(g.W), at(g.Uinv), at(g.Winv),
                          at(g.LinvT), b));
  // A_p(b:, :) = Q1_p(b:, :)·U_p⁻¹   (hh_recon.cxx:294-301)
  CQR_TRY(update_y_batched(c, g.d_bp, g.d_seg + g.nseg_tree + g.nseg_q1, g.nseg_y, g.ytiles, b, at(g.Y),
                           at(g.Uinv), A));
  // Tᵀ_p = −L_p⁻¹·W_pᵀ   (W = −T·Y1ᵀ)
  return gemm_f64_batched_prob(c, 1, 1, b, b, b, -1.0, bmat(at(g.LinvT), b, bb), SM_UNIT_UPPER, bmat(at(g.W), b, bb),
                               SM_UPPER, 0.0, bmat(Tt, b, 0, BM_PROB_T), SM_NONE, np, g.d_bp, 0, 0, b);
}

int run_all(cqr_ctx* c, cqr_batch* bt, double* A, double* Tt) {
  for (const Group& g : bt->groups) CQR_TRY(run_group(c, bt, g, A, Tt));
  return CQR_OK;
}

// FP32 inputs are staged in FP64 (the kernels compute in FP64).
int run_dtype(cqr_ctx* c, cqr_batch* bt, void* A, void* Tt, int dtype) {
  if (dtype == 0) return run_all(c, bt, static_cast<double*>(A), static_cast<double*>(Tt));
  double* sA = bt->stage;
  double* sT = bt->stage + bt->total_a;
  CQR_TRY(convert_f32_f64(c, A, sA, bt->total_a, 1));
  CQR_TRY(run_all(c, bt, sA, sT));
  CQR_TRY(convert_f32_f64(c, sA, A, bt->total_a, 0));
  return convert_f32_f64(c, sT, Tt, bt->total_t, 0);
}

void drop_graph(cqr_batch* bt) {
  if (bt->gexec) (void)hipGraphExecDestroy(bt->gexec);
  bt->gexec = nullptr;
  bt->g_A = bt->g_T = nullptr;
  bt->g_dtype = -1;
}

}  // namespace

extern "C" {

int cqr_batch_create(cqr_ctx* c, cqr_batch** out, int64_t nprob, const int64_t* m, const int64_t* n) {
  CQR_CHECK_CTX(c);
  if (!out || nprob < 1 || !m || !n) return set_err(c, CQR_ERR_INVALID, "batch_create: bad arguments");
  *out = nullptr;
  cqr_batch* bt = new cqr_batch();
  bt->nprob = nprob;
  bt->m.assign(m, m + nprob);
  bt->n.assign(n, n + nprob);
  bt->aoff.resize(nprob);
  bt->toff.resize(nprob);
  for (int64_t p = 0; p < nprob; ++p) {
    if (n[p] < 1 || n[p] > CQR_MAX_B || m[p] < n[p]) {
      delete bt;
      return set_err(c, n[p] > CQR_MAX_B ? CQR_ERR_UNSUPPORTED : CQR_ERR_INVALID,
                     "batch_create: problem %lld is %lldx%lld (needs m >= n, n <= %d)", (long long)p,
                     (long long)m[p], (long long)n[p], CQR_MAX_B);
    }
    bt->aoff[p] = bt->total_a;
    bt->toff[p] = bt->total_t;
    bt->total_a += m[p] * n[p];
    bt->total_t += n[p] * n[p];
  }
  for (int64_t p = 0; p < nprob; ++p) {
    Group* g = nullptr;
    for (Group& x : bt->groups)
      if (x.b == n[p]) g = &x;
    if (!g) {
      bt->groups.emplace_back();
      g = &bt->groups.back();
      g->b = (int)n[p];
    }
    g->probs.push_back((int)p);
  }
  for (Group& g : bt->groups) {
    const int s = plan_group(c, bt, g);
    if (s != CQR_OK) {
      for (Group& x : bt->groups) if (x.d_bp) (void)hipFree(x.d_bp);
      delete bt;
      return s;
    }
  }
  *out = bt;
  return CQR_OK;
}

int cqr_batch_destroy(cqr_ctx* c, cqr_batch* bt) {
  if (!bt) return CQR_OK;
  if (c) (void)hipStreamSynchronize(c->stream);
  drop_graph(bt);
  if (bt->cap) (void)hipStreamDestroy(bt->cap);
  for (Group& g : bt->groups) if (g.d_bp) (void)hipFree(g.d_bp);
  if (bt->ws) (void)hipFree(bt->ws);
  if (bt->stage) (void)hipFree(bt->stage);
  delete bt;
  return CQR_OK;
}

int cqr_batch_offsets(const cqr_batch* bt, int64_t* aoff, int64_t* toff, int64_t* total_a, int64_t* total_t) {
  if (!bt) return CQR_ERR_INVALID;
  if (aoff) std::memcpy(aoff, bt->aoff.data(), sizeof(int64_t) * bt->nprob);
  if (toff) std::memcpy(toff, bt->toff.data(), sizeof(int64_t) * bt->nprob);
  if (total_a) *total_a = bt->total_a;
  if (total_t) *total_t = bt->total_t;
  return CQR_OK;
}

int cqr_batch_groups(const cqr_batch* bt) { return bt ? (int)bt->groups.size() : 0; }

int cqr_batch_group_info(const cqr_batch* bt, int gi, int64_t* b, int64_t* nprob, int64_t* leaf_rows,
                         int64_t* leaves, int64_t* nodes) {
  if (!bt || gi < 0 || gi >= (int)bt->groups.size()) return CQR_ERR_INVALID;
  const Group& g = bt->groups[gi];
  if (b) *b = g.b;
  if (nprob) *nprob = (int64_t)g.probs.size();
  if (leaf_rows) *leaf_rows = g.max_leaf_rows;
  if (leaves) *leaves = g.total_leaves;
  if (nodes) *nodes = g.total_nodes;
  return CQR_OK;
}

int cqr_batch_set_leaf_rows(cqr_ctx* c, cqr_batch* bt, int gi, int64_t leaf_rows) {
  CQR_CHECK_CTX(c);
  if (!bt || gi < 0 || gi >= (int)bt->groups.size() || leaf_rows < 0)
    return set_err(c, CQR_ERR_INVALID, "batch_set_leaf_rows: bad arguments");
  CQR_HIP(c, hipStreamSynchronize(c->stream));
  drop_graph(bt);
  bt->groups[gi].leaf_rows = leaf_rows;
  return plan_group(c, bt, bt->groups[gi]);
}

int cqr_batch_qr(cqr_ctx* c, cqr_batch* bt, void* A, void* Tt, int dtype, int use_graph) {
  CQR_CHECK_CTX(c);
  if (!bt || !A || !Tt || (dtype != 0 && dtype != 1)) return set_err(c, CQR_ERR_INVALID, "batch_qr: bad arguments");
  CQR_TRY(ensure_batch_ws(c, bt));
  if (dtype == 1 && !bt->stage) {
    CQR_HIP(c, hipStreamSynchronize(c->stream));
    if (hipMalloc(&bt->stage, sizeof(double) * (bt->total_a + bt->total_t)) != hipSuccess) {
      (void)hipGetLastError();
      bt->stage = nullptr;
      return set_err(c, CQR_ERR_NOMEM, "batch FP64 staging");
    }
  }
  if (!use_graph) return run_dtype(c, bt, A, Tt, dtype);
  if (bt->gexec && bt->g_A == A && bt->g_T == Tt && bt->g_dtype == dtype && bt->g_stream == c->stream) {
    CQR_HIP(c, hipGraphLaunch(bt->gexec, c->stream));
    return CQR_OK;
  }
  // First call with these buffers: run eagerly (kernel attributes get set
  // outside the capture), then capture the same sequence for later calls.
  drop_graph(bt);
  CQR_TRY(run_dtype(c, bt, A, Tt, dtype));
  if (!bt->cap) CQR_HIP(c, hipStreamCreateWithFlags(&bt->cap, hipStreamNonBlocking));
  const int prof = c->prof;
  hipStream_t user = c->stream;
  c->prof = 0;
  c->stream = bt->cap;  // the launches below are recorded, not executed
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(bt->cap, hipStreamCaptureModeThreadLocal);
  int rc = e == hipSuccess ? run_dtype(c, bt, A, Tt, dtype) : CQR_ERR_DEVICE;
  if (e == hipSuccess) e = hipStreamEndCapture(bt->cap, &graph);
  c->stream = user;
  c->prof = prof;
  if (e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return set_err(c, CQR_ERR_DEVICE, "batch_qr: graph capture failed: %s", hipGetErrorString(e));
  }
  if (rc != CQR_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  const hipError_t ei = hipGraphInstantiate(&bt->gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    bt->gexec = nullptr;
    return set_err(c, CQR_ERR_DEVICE, "batch_qr: graph instantiation failed: %s", hipGetErrorString(ei));
  }
  bt->g_A = A;
  bt->g_T = Tt;
  bt->g_dtype = dtype;
  bt->g_stream = c->stream;
  return CQR_OK;
}


// The baseline the batch is measured against: hh_recon_qr per problem,
// one launch sequence after another.
int cqr_batch_qr_loop(cqr_ctx* c, cqr_batch* bt, double* A, double* Tt) {
  CQR_CHECK_CTX(c);
  if (!bt || !A || !Tt) return set_err(c, CQR_ERR_INVALID, "batch_qr_loop: bad arguments");
  size_t need = 0;
  for (int64_t p = 0; p < bt->nprob; ++p)
    need = std::max(need, panel_ws_bytes(bt->m[p], bt->n[p], 0, c->leaf_rows) + 8 * (size_t)bt->n[p] * bt->n[p] + 256);
  CQR_TRY(ensure_ws(c, need));
  char* base = static_cast<char*>(c->ws);
  for (int64_t p = 0; p < bt->nprob; ++p) {
    const int64_t mp = bt->m[p], np = bt->n[p];
    double* W = reinterpret_cast<double*>(base + need - 8 * np * np - 128);
    CQR_TRY(hh_recon_panel(c, A + bt->aoff[p], mp, mp, np, W, nullptr, nullptr, Tt + bt->toff[p], 0, base, nullptr));
  }
  return CQR_OK;
}

// ---------------------------------------------------------------------------
// Least squares for many requests at once: one batched factorization of all
// of them (the launches above, per width group), then each request's solve
// and FP64 refinement (lstsq_refine) from its factors.  Everything lives in
// the context's arena; no other allocation.
// ---------------------------------------------------------------------------
namespace {

struct LsqBatchPlan {
  cqr_batch bt;
  size_t oA = 0, oT = 0, oBp = 0, oG = 0, oR = 0, bytes = 0;
  std::vector<size_t> bp_off;     // per group, inside oBp
  size_t refine = 0;
};

int plan_lsq_batch(cqr_ctx* c, int64_t np, const cqr_request* rq, LsqBatchPlan& P) {
  cqr_batch& bt = P.bt;
  bt.nprob = np;
  bt.m.resize(np);
  bt.n.resize(np);
  bt.aoff.resize(np);
  bt.toff.resize(np);
  size_t refine = 0;
  for (int64_t p = 0; p < np; ++p) {
    const cqr_request& r = rq[p];
    if (!r.A || !r.B || !r.X || r.n < 1 || r.m < r.n || r.lda < r.m || r.ldb < r.m || r.ldx < r.n || r.nrhs < 1)
      return set_err(c, CQR_ERR_INVALID, "lstsq_batched: request %lld has bad arguments", (long long)p);
    if (r.n > CQR_MAX_B || r.nrhs > 16)
      return set_err(c, CQR_ERR_UNSUPPORTED, "lstsq_batched: request %lld needs n <= %d and nrhs <= 16",
                     (long long)p, CQR_MAX_B);
    bt.m[p] = r.m;
    bt.n[p] = r.n;
    bt.aoff[p] = bt.total_a;
    bt.toff[p] = bt.total_t;
    bt.total_a += r.m * r.n;
    bt.total_t += r.n * r.n;
    LsqWS w;
    refine = std::max(refine, refine_ws(r.m, r.n, r.nrhs, r.n, &w));
  }
  for (int64_t p = 0; p < np; ++p) {
    Group* g = nullptr;
    for (Group& x : bt.groups)
      if (x.b == bt.n[p]) g = &x;
    if (!g) {
      bt.groups.emplace_back();
      g = &bt.groups.back();
      g->b = (int)bt.n[p];
    }
    g->probs.push_back((int)p);
  }
  Sizer s;
  P.oA = s.add(8 * (size_t)bt.total_a);
  P.oT = s.add(8 * (size_t)bt.total_t);
  size_t gws = 0;
  for (Group& g : bt.groups) {
    CQR_TRY(plan_group(c, &bt, g, false));
    gws = std::max(gws, g.bytes);
    P.bp_off.push_back(s.add(g.htab.size()));
  }
  P.oG = s.add(gws);
  P.oR = s.add(refine);
  P.refine = refine;
  P.bytes = s.bytes;
  return CQR_OK;
}

}  // namespace

size_t cqr_lstsq_batched_workspace_bytes(int64_t nprob, const cqr_request* reqs) {
  if (nprob < 1 || !reqs) return 0;
  for (int64_t p = 0; p < nprob; ++p) {
    const cqr_request& r = reqs[p];
    if (r.n < 1 || r.n > CQR_MAX_B || r.m < r.n || r.nrhs < 1 || r.nrhs > 16) return 0;
  }
  cqr_ctx dummy;   // planning only: set_err needs somewhere to write
  LsqBatchPlan P;
  if (plan_lsq_batch(&dummy, nprob, reqs, P) != CQR_OK) return 0;
  return P.bytes;
}

int cqr_lstsq_batched(cqr_ctx* c, int64_t nprob, const cqr_request* reqs, int max_iter, double tol, int* status,
                      int* iters, double* last_update) {
  CQR_CHECK_CTX(c);
  if (nprob < 1 || !reqs || !status || max_iter < 0 || !(tol >= 0.0))
    return set_err(c, CQR_ERR_INVALID, "lstsq_batched: bad arguments");
  LsqBatchPlan P;
  CQR_TRY(plan_lsq_batch(c, nprob, reqs, P));
  CQR_TRY(ensure_ws(c, P.bytes));
  char* base = static_cast<char*>(c->ws);
  cqr_batch& bt = P.bt;
  bt.ws = base + P.oG;
  bt.ws_bytes = P.bytes - P.oG;
  double* Ap = reinterpret_cast<double*>(base + P.oA);
  double* Tp = reinterpret_cast<double*>(base + P.oT);
  for (size_t gi = 0; gi < bt.groups.size(); ++gi) {
    Group& g = bt.groups[gi];
    set_table(g, base + P.bp_off[gi]);
    // the tables outlive this call's last use of them (it ends in a sync)
    CQR_HIP(c, hipMemcpyAsync(g.d_bp, g.htab.data(), g.htab.size(), hipMemcpyHostToDevice,
                              c->stream));
  }
  for (int64_t p = 0; p < nprob; ++p)
    CQR_TRY(launch_copy2d(c, reqs[p].A, reqs[p].lda, Ap + bt.aoff[p], bt.m[p], bt.m[p], bt.n[p]));
  int rc = CQR_OK;
  for (Group& g : bt.groups)
    if ((rc = run_group(c, &bt, g, Ap, Tp)) != CQR_OK) break;
  for (int64_t p = 0; p < nprob; ++p) {
    const cqr_request& r = reqs[p];
    LsqRun run;
    LsqWS w;
    refine_ws(r.m, r.n, r.nrhs, r.n, &w);
    status[p] = rc != CQR_OK ? rc
                             : lstsq_refine(c, r.A, r.lda, r.m, r.n, r.B, r.ldb, r.nrhs, r.X, r.ldx, r.n,
                                            Ap + bt.aoff[p], r.m, Tp + bt.toff[p], CQR_PREC_F64, max_iter, tol, false,
                                            base + P.oR, w, run);
    if (iters) iters[p] = run.iters;
    if (last_update) last_update[p] = run.rel;
  }
  for (Group& g : bt.groups) g.d_bp = nullptr;   // arena memory, not owned
  bt.ws = nullptr;
  CQR_HIP(c, hipStreamSynchronize(c->stream));
  return rc;
}

// Leaf-height autotuner: for every group, time each candidate on a copy of
// A (two runs after a warm-up) and keep the fastest.
int cqr_batch_tune(cqr_ctx* c, cqr_batch* bt, const double* A, int reps) {
  CQR_CHECK_CTX(c);
  if (!bt || !A) return set_err(c, CQR_ERR_INVALID, "batch_tune: bad arguments");
  reps = std::max(reps, 1);
  double *scratchA = nullptr, *scratchT = nullptr;
  if (hipMalloc(&scratchA, sizeof(double) * bt->total_a) != hipSuccess ||
      hipMalloc(&scratchT, sizeof(double) * bt->total_t) != hipSuccess) {
    (void)hipGetLastError();
    if (scratchA) (void)hipFree(scratchA);
    return set_err(c, CQR_ERR_NOMEM, "batch_tune: scratch copy of A");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = CQR_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = CQR_ERR_DEVICE;
  drop_graph(bt);
  for (size_t gi = 0; gi < bt->groups.size() && rc == CQR_OK; ++gi) {
    Group& g = bt->groups[gi];
    const int64_t cap = leaf_capacity_rows(g.b);
    std::vector<int64_t> cand;
    for (int64_t r = cap; r >= std::max<int64_t>(128, g.b); r /= 2) cand.push_back(r);
    int64_t best = 0;
    float best_ms = 1e30f;
    for (int64_t rows : cand) {
      g.leaf_rows = rows;
      if ((rc = plan_group(c, bt, g)) != CQR_OK || (rc = ensure_batch_ws(c, bt)) != CQR_OK) break;
      float tot = 0.0f;
      for (int r = 0; r <= reps && rc == CQR_OK; ++r) {
        if (hipMemcpyAsync(scratchA, A, sizeof(double) * bt->total_a, hipMemcpyDeviceToDevice, c->stream) !=
            hipSuccess) { rc = CQR_ERR_DEVICE; break; }
        (void)hipEventRecord(e0, c->stream);
        rc = run_group(c, bt, g, scratchA, scratchT);
        (void)hipEventRecord(e1, c->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r > 0) tot += ms;  // the first run warms up
      }
      if (rc == CQR_OK && tot < best_ms) {
        best_ms = tot;
        best = rows;
      }
    }
    if (rc == CQR_OK) {
      g.leaf_rows = best;
      rc = plan_group(c, bt, g);
    }
  }
  if (rc == CQR_OK) rc = ensure_batch_ws(c, bt);
  (void)hipStreamSynchronize(c->stream);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(scratchA);
  (void)hipFree(scratchT);
  if (rc == CQR_ERR_DEVICE) return set_err(c, rc, "batch_tune: device error");
  return rc;
}

}  // extern "C"
