// Least-squares request server with continuous batching on one GPU.
//
// Requests (A, B resident in device memory, X written back to device
// memory) enter a FIFO queue.  A fixed set of slots serves them; each slot
// owns a library context (its own HIP stream and workspace arena) and a host
// thread that drives it.  A slot takes the oldest queued request as soon as
// it is idle and the request's HBM footprint fits the remaining budget, so
// requests that arrive while others run join immediately and their kernels
// overlap on the device.  Every result depends only on its own request
// (each slot factors into its own arena; all reductions have a fixed
// order), so a request gives the same bits alone and in any mix.
//
// Two accuracy tiers: FP64 (trailing update in FP64, refined to 1e-12) and
// mixed (refined with FP64 residuals to 1e-10; FP32 trailing update from
// kMixedMinCols columns on, where it is faster).
//
// Requests of one panel (n ≤ 128, so no trailing update to make mixed) are
// batched: a slot that takes such a request also takes the queued requests
// of the same tier and width behind it (up to max_batch, while the budget
// allows) and runs them through cqr_lstsq_batched, one batched
// factorization and then each solve.  A batched request's X depends only on
// that request, so it gets the same bits in a batch of any size.
// The serving loop has no counterpart in the reference.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/candmc_qr.h"

namespace {

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Job {
  int64_t id;
  cqr_request req;
  size_t bytes;     // footprint: batched (one-panel requests) or single solve
  bool batch;       // goes through the batched path
  cqr_result res;
};

bool one_panel(const cqr_request& r) { return r.n <= 128 && r.nrhs <= 16; }

}  // namespace

struct cqr_server {
  int device = 0;
  size_t budget = 0, in_use = 0;   // in_use: Σ the slots' resident arenas
  std::vector<size_t> resident;    // per slot: its context's arena bytes
  int waiters = 0;                 // threads inside cqr_server_wait
  int max_batch = 32;              // one-panel requests per batch (1: no batching)
  int64_t next_id = 0;
  bool stop = false;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job> queue;
  std::unordered_map<int64_t, cqr_result> done;   // finished, not yet collected
  std::unordered_set<int64_t> live;               // submitted, not yet collected
  std::vector<int64_t> done_order;
  std::vector<cqr_ctx*> ctxs;
  std::vector<hipStream_t> streams;
  std::vector<std::thread> workers;
};

namespace {

// Bytes a context's arena holds after it has grown for a request of
// `bytes` (the arena keeps an eighth of headroom when it grows).
size_t arena_after(const cqr_ctx* c, size_t bytes) {
  const size_t have = cqr_workspace_bytes(c);
  return have >= bytes ? have : bytes + bytes / 8;
}

// Runs a batch of one-panel requests of one tier; fills their results.
void run_batch(cqr_ctx* c, std::vector<Job>& jobs, int slot) {
  const int64_t t0 = now_ns();
  const size_t k = jobs.size();
  std::vector<cqr_request> rq(k);
  std::vector<int> st(k, CQR_OK), it(k, 0);
  std::vector<double> upd(k, 0.0);
  size_t bytes = 0;
  for (size_t j = 0; j < k; ++j) {
    rq[j] = jobs[j].req;
    bytes += jobs[j].bytes;
  }
  const bool mixed = rq[0].tier == CQR_TIER_MIXED;
  int rc = cqr_reserve(c, bytes);
  if (rc == CQR_OK)
    rc = cqr_lstsq_batched(c, (int64_t)k, rq.data(), mixed ? 30 : 4, mixed ? 1e-10 : 1e-12, st.data(), it.data(),
                           upd.data());
  if (rc == CQR_OK) rc = cqr_sync(c);
  const int64_t t1 = now_ns();
  for (size_t j = 0; j < k; ++j) {
    cqr_result& r = jobs[j].res;
    r.t_start_ns = t0;
    r.t_end_ns = t1;
    r.status = rc != CQR_OK ? rc : st[j];
    r.iters = it[j];
    r.last_update = upd[j];
    r.prec = rc != CQR_OK ? -1 : CQR_PREC_F64;   // one panel: factored in FP64
    r.slot = slot;
    r.batch = (int)k;
  }
}

// The mixed tier factors with the FP32 trailing update only from this many
// columns on: below it the panels (FP64 in every tier) dominate and the
// FP32 update only adds refinement steps (measured: at n ≤ 256 the mixed
// factorization served 15 % fewer solves/s; at 65536×2048 it factored 5 %
// faster), so narrower mixed requests are factored in FP64 and refined to
// the mixed tier's tolerance.
constexpr int64_t kMixedMinCols = 1024;

// Runs one request through cqr_lstsq; fills its result.
void run_single(cqr_ctx* c, Job& job, int slot) {
  const cqr_request& r = job.req;
  job.res.t_start_ns = now_ns();
  const bool mixed = r.tier == CQR_TIER_MIXED;
  const int prec = mixed && r.n >= kMixedMinCols ? CQR_PREC_F32 : CQR_PREC_F64;
  int iters = 0;
  double upd = 0.0;
  // the arena grows here, to what admission charged, not inside the solve
  int st = cqr_reserve(c, job.bytes);
  if (st == CQR_OK)
    st = cqr_lstsq(c, r.A, r.lda, r.m, r.n, r.B, r.ldb, r.nrhs, r.X, r.ldx, 128, prec, mixed ? 30 : 4,
                   mixed ? 1e-10 : 1e-12, &iters, &upd);
  if (st == CQR_OK || st == CQR_ERR_NOCONVERGE) {
    const int ss = cqr_sync(c);
    if (ss != CQR_OK) st = ss;
  }
  job.res.t_end_ns = now_ns();
  job.res.status = st;
  job.res.iters = iters;
  job.res.last_update = upd;
  job.res.prec = cqr_lstsq_last_prec(c);
  job.res.slot = slot;
  job.res.batch = 1;
}

void serve(cqr_server* s, int slot) {
  cqr_ctx* c = s->ctxs[slot];
  (void)hipSetDevice(s->device);
  for (;;) {
    std::vector<Job> jobs;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop) return;
        if (!s->queue.empty()) {
          // FIFO: the head is admitted when this slot's arena, grown for it,
          // keeps the sum of all resident arenas within the budget; a
          // one-panel head takes the one-panel requests of its tier queued
          // behind it along, as far as the budget and max_batch allow
          const Job& head = s->queue.front();
          const size_t after = arena_after(c, head.bytes);
          if (s->in_use - s->resident[slot] + after <= s->budget) {
            size_t sum = head.bytes, charge = after;
            std::vector<size_t> pick{0};
            if (head.batch)
              for (size_t q = 1; q < s->queue.size() && (int)pick.size() < s->max_batch; ++q) {
                const Job& x = s->queue[q];
                // one width: a batch is then one launch sequence (requests of
                // other widths stay queued for the other slots, which would
                // otherwise idle while this one ran the widths in turn)
                if (!x.batch || x.req.tier != head.req.tier || x.req.n != head.req.n) continue;
                const size_t a = arena_after(c, sum + x.bytes);
                if (s->in_use - s->resident[slot] + a > s->budget) break;
                sum += x.bytes;
                charge = a;
                pick.push_back(q);
              }
            for (size_t q : pick) jobs.push_back(s->queue[q]);
            for (size_t i = pick.size(); i-- > 0;) s->queue.erase(s->queue.begin() + (long)pick[i]);
            s->in_use += charge - s->resident[slot];
            s->resident[slot] = charge;
            break;
          }
          if (s->resident[slot] > 0) {
            // idle with an arena the head cannot share: give it back
            const size_t mine = s->resident[slot];
            lk.unlock();
            (void)cqr_trim(c);
            lk.lock();
            s->in_use -= mine;
            s->resident[slot] = 0;
            s->cv_work.notify_all();
            continue;
          }
        }
        s->cv_work.wait(lk);
      }
    }
    if (jobs[0].batch) run_batch(c, jobs, slot);
    else run_single(c, jobs[0], slot);
    const size_t held = cqr_workspace_bytes(c);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      // the arena stays resident with the slot (charged until released)
      s->in_use += held - s->resident[slot];
      s->resident[slot] = held;
      for (Job& job : jobs) {
        job.res.bytes = (int64_t)held;
        s->done[job.id] = job.res;
        s->done_order.push_back(job.id);
      }
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
  }
}

}  // namespace

extern "C" {

int cqr_server_create(cqr_server** out, int device, int nslots, size_t hbm_budget) {
  if (!out || nslots < 1 || nslots > 64 || hbm_budget == 0) return CQR_ERR_INVALID;
  *out = nullptr;
  cqr_server* s = new cqr_server();
  s->device = device;
  s->budget = hbm_budget;
  for (int i = 0; i < nslots; ++i) {
    cqr_ctx* c = nullptr;
    int st = cqr_create(&c, device);
    hipStream_t stream = nullptr;
    if (st == CQR_OK && hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) st = CQR_ERR_DEVICE;
    if (st == CQR_OK) st = cqr_set_stream(c, stream);
    if (st != CQR_OK) {
      if (c) cqr_destroy(c);
      if (stream) (void)hipStreamDestroy(stream);
      for (size_t k = 0; k < s->ctxs.size(); ++k) {
        cqr_destroy(s->ctxs[k]);
        (void)hipStreamDestroy(s->streams[k]);
      }
      delete s;
      return st;
    }
    s->ctxs.push_back(c);
    s->streams.push_back(stream);
  }
  s->resident.assign(nslots, 0);
  for (int i = 0; i < nslots; ++i) s->workers.emplace_back(serve, s, i);
  *out = s;
  return CQR_OK;
}

int cqr_server_destroy(cqr_server* s) {
  if (!s) return CQR_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (std::thread& t : s->workers) t.join();
  {
    // requests that never started finish as cancelled, so a thread waiting
    // for one returns; then wait for every waiter to leave
    std::unique_lock<std::mutex> lk(s->mu);
    for (Job& j : s->queue) {
      j.res.status = CQR_ERR_CANCELLED;
      j.res.prec = -1;
      j.res.slot = -1;
      j.res.batch = 0;
      s->done[j.id] = j.res;
      s->done_order.push_back(j.id);
    }
    s->queue.clear();
    s->cv_done.notify_all();
    s->cv_done.wait(lk, [&] { return s->waiters == 0; });
  }
  for (size_t k = 0; k < s->ctxs.size(); ++k) {
    cqr_destroy(s->ctxs[k]);
    (void)hipStreamDestroy(s->streams[k]);
  }
  delete s;
  return CQR_OK;
}

int64_t cqr_server_submit(cqr_server* s, const cqr_request* r) {
  if (!s || !r) return -CQR_ERR_INVALID;
  if (!r->A || !r->B || !r->X || r->m < r->n || r->n < 1 || r->nrhs < 1 || r->nrhs > 16 || r->lda < r->m ||
      r->ldb < r->m || r->ldx < r->n || (r->tier != CQR_TIER_F64 && r->tier != CQR_TIER_MIXED))
    return -CQR_ERR_INVALID;
  bool batch;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    batch = s->max_batch > 1 && one_panel(*r);
  }
  const size_t bytes =
      batch ? cqr_lstsq_batched_workspace_bytes(1, r) : cqr_lstsq_workspace_bytes(r->m, r->n, r->nrhs, 128);
  // can never be admitted: its arena (with headroom) exceeds the budget
  if (bytes == 0 || bytes + bytes / 8 > s->budget) return -CQR_ERR_NOMEM;
  Job j;
  j.req = *r;
  j.bytes = bytes;
  j.batch = batch;
  j.res = cqr_result{};
  j.res.t_submit_ns = now_ns();
  j.res.bytes = (int64_t)bytes;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    j.id = s->next_id++;
    j.res.id = j.id;
    s->queue.push_back(j);
    s->live.insert(j.id);
  }
  s->cv_work.notify_all();
  return j.id;
}

int cqr_server_wait(cqr_server* s, int64_t id, cqr_result* out) {
  if (!s || id < 0) return CQR_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  // an id never submitted, or already collected by a wait or a drain, has
  // nothing left to wait for
  if (!s->live.count(id)) return CQR_ERR_INVALID;
  ++s->waiters;
  s->cv_done.wait(lk, [&] { return s->done.count(id) > 0 || !s->live.count(id); });
  --s->waiters;
  if (s->waiters == 0) s->cv_done.notify_all();   // a destroy may wait for this
  if (!s->done.count(id)) return CQR_ERR_INVALID;  // a concurrent drain took it
  if (out) *out = s->done[id];
  s->done.erase(id);
  s->live.erase(id);
  for (size_t k = 0; k < s->done_order.size(); ++k)
    if (s->done_order[k] == id) {
      s->done_order.erase(s->done_order.begin() + (long)k);
      break;
    }
  return CQR_OK;
}

int cqr_server_drain(cqr_server* s, cqr_result* out, int max, int* count) {
  if (!s || (max > 0 && !out) || !count) return CQR_ERR_INVALID;
  int k = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<int64_t> keep;
    for (int64_t id : s->done_order) {
      if (k < max) {
        out[k++] = s->done[id];
        s->done.erase(id);
        s->live.erase(id);
      } else {
        keep.push_back(id);
      }
    }
    s->done_order.swap(keep);
  }
  if (k) s->cv_done.notify_all();  // waiters on a drained id return
  *count = k;
  return CQR_OK;
}

int cqr_server_set_batching(cqr_server* s, int max_batch) {
  if (!s || max_batch < 1 || max_batch > 4096) return CQR_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  s->max_batch = max_batch;
  return CQR_OK;
}

int cqr_server_pending(cqr_server* s, int64_t* queued, size_t* bytes_in_use) {
  if (!s) return CQR_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (queued) *queued = (int64_t)s->queue.size();
  if (bytes_in_use) *bytes_in_use = s->in_use;   // resident arenas
  return CQR_OK;
}

}  // extern "C"
