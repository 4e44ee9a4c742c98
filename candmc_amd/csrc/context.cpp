// Context, errors, workspace arena and TSQR planning (host only).
#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <tuple>

#include "cqr_internal.h"

namespace cqr {

int set_err(cqr_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

int check_launch(cqr_ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(c, CQR_ERR_DEVICE, "%s launch: %s", what, hipGetErrorString(e));
  return CQR_OK;
}

// Launches whose workgroups wait on other workgroups of the same launch
// (the TSQR tree, construct_Q1) make progress alone: each XCD dispatches a
// launch's blocks in order, and a block only waits on blocks of lower
// index.  Two such launches at once on different streams have no common
// order: each can fill an XCD's CUs with blocks that wait on producers the
// other keeps from being dispatched, until their poll bound runs out (seen
// with 16 server slots: seconds per request).  So such launches run
// together only while all of their blocks fit on the device at once: each
// launch's need is its blocks per XCD (they are dealt round-robin to the
// eight XCDs) over the blocks a CU holds (the occupancy API), against the
// CUs of one XCD.  A launch that would overflow it first waits for the
// oldest running ones to finish (a stream wait on their completion
// events; finished ones are dropped on the host); one that needs the whole
// device runs alone.  Other launches (GEMMs, the leaf QR) are not counted:
// their blocks always finish.  A stream being captured into a graph is not
// gated (a capture cannot wait on an event recorded outside it).
namespace {
constexpr int kGateDevices = 64, kGateEvents = 64;
struct SpinGate {
  std::mutex mu;
  struct Running {
    hipEvent_t ev;
    int need;
  };
  std::deque<Running> running;   // launch order
  std::vector<hipEvent_t> pool;  // free events (all created up front)
  int used = 0, cap = 0;
  bool ready = false;
};
SpinGate& spin_gate(int device) {
  static SpinGate gates[kGateDevices];
  return gates[device];
}
// blocks of `fn` one CU holds (cached)
int blocks_per_cu(const void* fn, int threads, size_t lds) {
  static std::mutex mu;
  static std::map<std::tuple<const void*, int, size_t>, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  const auto key = std::make_tuple(fn, threads, lds);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, threads, lds) != hipSuccess || n < 1) {
    (void)hipGetLastError();
    n = 1;
  }
  cache[key] = n;
  return n;
}
}  // namespace

int spin_gate_init(cqr_ctx* c) {
  if (c->device < 0 || c->device >= kGateDevices) return CQR_OK;
  SpinGate& g = spin_gate(c->device);
  std::lock_guard<std::mutex> lk(g.mu);
  if (g.ready) return CQR_OK;
  for (int k = 0; k < kGateEvents; ++k) {
    hipEvent_t e = nullptr;
    CQR_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    g.pool.push_back(e);
  }
  g.cap = std::max(1, c->num_cu / 8);
  g.ready = true;
  return CQR_OK;
}

SpinScope::SpinScope(cqr_ctx* ctx, const void* fn, int64_t blocks, int threads, size_t lds) : c(ctx) {
  if (c->device < 0 || c->device >= kGateDevices) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    (void)hipGetLastError();
    return;
  }
  SpinGate& g = spin_gate(c->device);
  g.mu.lock();
  if (!g.ready) {
    g.mu.unlock();
    return;
  }
  held = true;
  const int64_t per_xcd = (blocks + 7) / 8, per_cu = blocks_per_cu(fn, threads, lds);
  need = (int)std::min<int64_t>(g.cap, (per_xcd + per_cu - 1) / per_cu);
  // finished launches give their room back
  while (!g.running.empty() && hipEventQuery(g.running.front().ev) == hipSuccess) {
    g.used -= g.running.front().need;
    g.pool.push_back(g.running.front().ev);
    g.running.pop_front();
  }
  (void)hipGetLastError();   // hipEventQuery's hipErrorNotReady
  // then wait for the oldest running ones until this launch fits
  while (!g.running.empty() && (g.used + need > g.cap || g.pool.empty())) {
    if (hipStreamWaitEvent(c->stream, g.running.front().ev, 0) != hipSuccess)
      status_ = set_err(c, CQR_ERR_DEVICE, "cannot order a waiting launch after a running one");
    g.used -= g.running.front().need;
    g.pool.push_back(g.running.front().ev);
    g.running.pop_front();
  }
}

SpinScope::~SpinScope() {
  if (!held) return;
  SpinGate& g = spin_gate(c->device);
  hipEvent_t e = g.pool.back();
  if (hipEventRecord(e, c->stream) == hipSuccess) {
    g.pool.pop_back();
    g.running.push_back(SpinGate::Running{e, need});
    g.used += need;
  } else {
    (void)hipGetLastError();
  }
  g.mu.unlock();
}

int set_max_lds(cqr_ctx* c, const void* fn, int bytes) {
  static std::mutex mu;
  static std::set<std::tuple<const void*, int, int>> done;
  std::lock_guard<std::mutex> lk(mu);
  const auto key = std::make_tuple(fn, c->device, bytes);
  if (done.count(key)) return CQR_OK;
  CQR_HIP(c, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  done.insert(key);
  return CQR_OK;
}

static hipEvent_t pool_get(cqr_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return e;
}

hipEvent_t prof_begin(cqr_ctx* c) {
  if (!c->prof) return nullptr;
  hipEvent_t a = pool_get(c);
  if (a) (void)hipEventRecord(a, c->stream);
  return a;
}

void prof_end(cqr_ctx* c, hipEvent_t a, int kind, double flops, double bytes) {
  if (!c->prof || !a) return;
  hipEvent_t b = pool_get(c);
  if (!b) {
    c->pool.push_back(a);  // no pair: the start event goes back unused
    return;
  }
  (void)hipEventRecord(b, c->stream);
  c->recs.push_back({a, b, kind, flops, bytes});
}

static int prof_drain(cqr_ctx* c) {
  if (c->recs.empty()) return CQR_OK;
  CQR_HIP(c, hipStreamSynchronize(c->stream));
  for (const ProfRec& r : c->recs) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      c->acc_ms[r.kind] += ms;
      c->acc_flops[r.kind] += r.flops;
      c->acc_bytes[r.kind] += r.bytes;
      c->acc_n[r.kind] += 1;
    } else {
      (void)hipGetLastError();
    }
    c->pool.push_back(r.a);
    c->pool.push_back(r.b);
  }
  c->recs.clear();
  return CQR_OK;
}

// The lookahead's second stream, created on first use: a context that never
// factors with lookahead (a served solve of a few panels) holds only its own
// stream.  HIP deals the streams of a process round-robin onto a few
// hardware queues (four by default), so every extra stream is one more that
// can land on the queue of another context's stream and serialise with it.
int ensure_side(cqr_ctx* c) {
  if (c->side) return CQR_OK;
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) {
    c->side = nullptr;
    (void)hipGetLastError();
    return set_err(c, CQR_ERR_DEVICE, "cannot create the lookahead stream");
  }
  return CQR_OK;
}

int ensure_ws(cqr_ctx* c, size_t bytes) {
  if (bytes <= c->ws_bytes) return CQR_OK;
  // Growing replaces the arena: wait for work that may still use it.
  if (c->ws) {
    CQR_HIP(c, hipStreamSynchronize(c->stream));
    if (c->side) CQR_HIP(c, hipStreamSynchronize(c->side));
    CQR_HIP(c, hipFree(c->ws));
    c->ws = nullptr;
    c->ws_bytes = 0;
  }
  size_t want = bytes + bytes / 8;  // headroom against small regrowths
  hipError_t e = hipMalloc(&c->ws, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    c->ws = nullptr;
    return set_err(c, CQR_ERR_NOMEM, "workspace of %zu bytes: %s", want, hipGetErrorString(e));
  }
  c->ws_bytes = want;
  return CQR_OK;
}

int make_plan(int64_t m, int64_t b, int64_t nleaves_hint, int64_t leaf_rows, TsqrPlan& p) {
  if (m < b || b < 1) return CQR_ERR_INVALID;
  if (b > CQR_MAX_B) return CQR_ERR_UNSUPPORTED;
  const int64_t cap = leaf_capacity_rows(b);
  int64_t rows = leaf_rows > 0 ? std::min(leaf_rows, cap) : cap;
  rows = std::max(rows, b);
  int64_t L = (m + rows - 1) / rows;                 // fewest leaves that fit
  L = std::max<int64_t>(L, std::max<int64_t>(nleaves_hint, 1));
  L = std::min<int64_t>(L, m / b);                   // every leaf keeps ≥ b rows
  L = std::max<int64_t>(L, 1);
  if ((m + L - 1) / L > cap) return CQR_ERR_UNSUPPORTED;
  p.m = m;
  p.b = b;
  p.nleaves = L;
  p.max_leaf_rows = (m + L - 1) / L;
  p.leaf_off.resize(L + 1);
  for (int64_t l = 0; l <= L; ++l) p.leaf_off[l] = (l * m) / L;
  p.levels.clear();
  p.level_node0.clear();
  p.nnodes = 0;
  for (int64_t n = L; n > 1; n = n / 2 + n % 2) {    // bitree_tsqr.cxx:172
    std::vector<std::pair<int, int>> lv;
    for (int64_t r = 0; r < n / 2; ++r) lv.push_back({(int)r, (int)(r + (n + 1) / 2)});
    p.level_node0.push_back(p.nnodes);
    p.nnodes += (int64_t)lv.size();
    p.levels.push_back(std::move(lv));
  }
  return CQR_OK;
}

}  // namespace cqr

using namespace cqr;

extern "C" {

int cqr_create(cqr_ctx** ctx, int device) {
  if (!ctx) return CQR_ERR_INVALID;
  *ctx = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    (void)hipGetLastError();
    return CQR_ERR_DEVICE;
  }
  if (hipSetDevice(device) != hipSuccess) return CQR_ERR_DEVICE;
  cqr_ctx* c = new cqr_ctx();
  c->device = device;
  if (hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      c->num_cu < 1)
    c->num_cu = 256;
  if (hipMalloc(&c->d_flag, sizeof(int)) != hipSuccess ||
      hipMalloc(&c->d_cnt, sizeof(uint64_t) * CQR_MAX_TREE_NODES * 2) != hipSuccess ||
      hipMalloc(&c->d_terr, 2 * sizeof(int)) != hipSuccess ||   // [0] timeout flag, [1] poll bound
      hipMalloc(&c->d_qflags, sizeof(uint64_t) * CQR_MAX_TREE_NODES * 8) != hipSuccess ||
      hipMemset(c->d_qflags, 0, sizeof(uint64_t) * CQR_MAX_TREE_NODES * 8) != hipSuccess ||
      hipMalloc(&c->d_sflags, sizeof(uint64_t) * CQR_TRSM_MAX_BLOCKS) != hipSuccess ||
      hipHostMalloc(&c->h_pin, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_host, hipEventDisableTiming) != hipSuccess ||
      hipMemset(c->d_terr, 0, 2 * sizeof(int)) != hipSuccess ||
      hipMemset(c->d_cnt, 0, sizeof(uint64_t) * CQR_MAX_TREE_NODES * 2) != hipSuccess ||
      hipMemset(c->d_flag, 0, sizeof(int)) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_gate, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    if (c->d_flag) (void)hipFree(c->d_flag);
    if (c->d_cnt) (void)hipFree(c->d_cnt);
    if (c->d_terr) (void)hipFree(c->d_terr);
    if (c->d_qflags) (void)hipFree(c->d_qflags);
    if (c->d_sflags) (void)hipFree(c->d_sflags);
    if (c->h_pin) (void)hipHostFree(c->h_pin);
    if (c->ev_host) (void)hipEventDestroy(c->ev_host);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_gate) (void)hipEventDestroy(c->ev_gate);
    delete c;
    return CQR_ERR_NOMEM;
  }
  if (cqr::spin_gate_init(c) != CQR_OK || cqr::noalloc_setup(c) != CQR_OK) {
    cqr_destroy(c);
    return CQR_ERR_DEVICE;
  }
  *ctx = c;
  return CQR_OK;
}

int cqr_destroy(cqr_ctx* c) {
  if (!c) return CQR_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->side) (void)hipStreamSynchronize(c->side);
  for (const ProfRec& r : c->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
  if (c->ws) (void)hipFree(c->ws);
  if (c->d_flag) (void)hipFree(c->d_flag);
  if (c->d_cnt) (void)hipFree(c->d_cnt);
  if (c->d_terr) (void)hipFree(c->d_terr);
  if (c->d_qflags) (void)hipFree(c->d_qflags);
  if (c->d_sflags) (void)hipFree(c->d_sflags);
  if (c->h_pin) (void)hipHostFree(c->h_pin);
  if (c->ev_host) (void)hipEventDestroy(c->ev_host);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_gate) (void)hipEventDestroy(c->ev_gate);
  delete c;
  return CQR_OK;
}

int cqr_set_stream(cqr_ctx* c, void* s) {
  if (!c) return CQR_ERR_INVALID;
  c->stream = static_cast<hipStream_t>(s);
  return CQR_OK;
}

void* cqr_get_stream(const cqr_ctx* c) { return c ? (void*)c->stream : nullptr; }

int cqr_sync(cqr_ctx* c) {
  if (!c) return CQR_ERR_INVALID;
  CQR_HIP(c, hipStreamSynchronize(c->stream));
  // a pipelined tree node that gave up waiting for a child leaves its flag
  int terr = 0;
  CQR_HIP(c, hipMemcpy(&terr, c->d_terr, sizeof(int), hipMemcpyDeviceToHost));
  if (terr) {
    CQR_HIP(c, hipMemset(c->d_terr, 0, sizeof(int)));
    return set_err(c, CQR_ERR_DEVICE, "TSQR tree node timed out waiting for a child");
  }
  return CQR_OK;
}

int cqr_set_wait_limit(cqr_ctx* c, int polls) {
  CQR_CHECK_CTX(c);
  if (polls < 0) return set_err(c, CQR_ERR_INVALID, "set_wait_limit: polls must be >= 0");
  CQR_HIP(c, hipStreamSynchronize(c->stream));
  CQR_HIP(c, hipMemcpy(c->d_terr + 1, &polls, sizeof(int), hipMemcpyHostToDevice));
  return CQR_OK;
}

int cqr_set_lookahead(cqr_ctx* c, int on) {
  if (!c) return CQR_ERR_INVALID;
  c->lookahead = on ? 1 : 0;
  return CQR_OK;
}

int cqr_reserve(cqr_ctx* c, size_t bytes) {
  if (!c) return CQR_ERR_INVALID;
  (void)hipSetDevice(c->device);
  return ensure_ws(c, bytes);
}

int cqr_trim(cqr_ctx* c) {
  if (!c) return CQR_ERR_INVALID;
  (void)hipSetDevice(c->device);
  if (!c->ws) return CQR_OK;
  CQR_HIP(c, hipStreamSynchronize(c->stream));
  if (c->side) CQR_HIP(c, hipStreamSynchronize(c->side));
  CQR_HIP(c, hipFree(c->ws));
  c->ws = nullptr;
  c->ws_bytes = 0;
  return CQR_OK;
}

size_t cqr_workspace_bytes(const cqr_ctx* c) { return c ? c->ws_bytes : 0; }

int cqr_set_check_finite(cqr_ctx* c, int on) {
  if (!c) return CQR_ERR_INVALID;
  c->check_finite = on ? 1 : 0;
  return CQR_OK;
}

int cqr_set_leaf_rows(cqr_ctx* c, int64_t rows) {
  if (!c || rows < 0) return CQR_ERR_INVALID;
  c->leaf_rows = rows;
  return CQR_OK;
}

int cqr_profile_enable(cqr_ctx* c, int on) {
  if (!c) return CQR_ERR_INVALID;
  c->prof = on ? 1 : 0;
  return CQR_OK;
}

int cqr_profile_read(cqr_ctx* c, int kind, double* ms, double* flops, double* bytes, int64_t* launches) {
  if (!c || kind < 0 || kind >= PK_NKINDS) return CQR_ERR_INVALID;
  CQR_TRY(prof_drain(c));
  if (ms) *ms = c->acc_ms[kind];
  if (flops) *flops = c->acc_flops[kind];
  if (bytes) *bytes = c->acc_bytes[kind];
  if (launches) *launches = c->acc_n[kind];
  return CQR_OK;
}

int cqr_profile_reset(cqr_ctx* c) {
  if (!c) return CQR_ERR_INVALID;
  CQR_TRY(prof_drain(c));
  for (int k = 0; k < PK_NKINDS; ++k) {
    c->acc_ms[k] = c->acc_flops[k] = c->acc_bytes[k] = 0.0;
    c->acc_n[k] = 0;
  }
  return CQR_OK;
}

const char* cqr_status_string(int s) {
  switch (s) {
    case CQR_OK: return "ok";
    case CQR_ERR_INVALID: return "invalid argument";
    case CQR_ERR_UNSUPPORTED: return "unsupported shape";
    case CQR_ERR_NOMEM: return "out of device memory";
    case CQR_ERR_DEVICE: return "HIP runtime error";
    case CQR_ERR_NONFINITE: return "non-finite input or result";
    case CQR_ERR_NOCONVERGE: return "refinement did not reach the tolerance";
    case CQR_ERR_CANCELLED: return "request cancelled";
    default: return "unknown status";
  }
}

const char* cqr_last_error(const cqr_ctx* c) { return c ? c->err.c_str() : ""; }

int cqr_plan_tsqr(int64_t m, int64_t b, int64_t nleaves_hint, int64_t leaf_rows, int64_t* nleaves,
                  int64_t* max_leaf_rows, int64_t* nlevels) {
  TsqrPlan p;
  const int s = make_plan(m, b, nleaves_hint, leaf_rows, p);
  if (s != CQR_OK) return s;
  if (nleaves) *nleaves = p.nleaves;
  if (max_leaf_rows) *max_leaf_rows = p.max_leaf_rows;
  if (nlevels) *nlevels = (int64_t)p.levels.size();
  return CQR_OK;
}

}  // extern "C"
