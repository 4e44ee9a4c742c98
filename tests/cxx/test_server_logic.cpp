// Host-only test of the request server's queue and collection logic
// (candmc_amd/csrc/server.cpp): the GPU calls it makes are replaced by the
// stand-ins below, so each "solve" is a short sleep.  Checked: a result is
// collected exactly once; a wait on an id never submitted, or already
// collected by a wait or a drain, returns CQR_ERR_INVALID instead of
// blocking; two waiters on one id, or a waiter racing a drain, end with
// exactly one collector; every request of a burst through two slots is
// served; the slots' resident arenas are what admission charges and never
// exceed the budget; destroying a server with requests queued cancels them
// (a waiter returns) and returns once the running one ends; one-panel
// requests of one width are batched up to max_batch, within the budget, and
// wider ones never are.  Built with g++ and run by tests/test_server_logic.py.
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "candmc_qr.h"

static std::atomic<int> solve_ms{5};

extern "C" {
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = reinterpret_cast<hipStream_t>(0x1);
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
// a stand-in context is one size_t: its arena bytes
int cqr_create(cqr_ctx** c, int) {
  *c = static_cast<cqr_ctx*>(calloc(1, sizeof(size_t)));
  return CQR_OK;
}
static std::atomic<size_t> arena_now{0}, arena_peak{0};  // Σ of the stand-in arenas
static void arena_set(cqr_ctx* c, size_t v) {
  size_t* p = reinterpret_cast<size_t*>(c);
  arena_now += v - *p;
  *p = v;
  size_t cur = arena_now.load(), pk = arena_peak.load();
  while (cur > pk && !arena_peak.compare_exchange_weak(pk, cur)) {}
}
int cqr_reserve(cqr_ctx* c, size_t bytes) {
  if (*reinterpret_cast<size_t*>(c) < bytes) arena_set(c, bytes + bytes / 8);
  return CQR_OK;
}
int cqr_trim(cqr_ctx* c) {
  arena_set(c, 0);
  return CQR_OK;
}
size_t cqr_workspace_bytes(const cqr_ctx* c) { return *reinterpret_cast<const size_t*>(c); }
int cqr_lstsq_last_prec(const cqr_ctx*) { return CQR_PREC_F64; }
int cqr_destroy(cqr_ctx* c) {
  arena_set(c, 0);
  free(c);
  return CQR_OK;
}
int cqr_set_stream(cqr_ctx*, void*) { return CQR_OK; }
int cqr_sync(cqr_ctx*) { return CQR_OK; }
int cqr_lstsq(cqr_ctx*, const double*, int64_t, int64_t, int64_t, const double*, int64_t, int64_t, double*, int64_t,
              int64_t, int, int, double, int* iters, double*) {
  std::this_thread::sleep_for(std::chrono::milliseconds(solve_ms.load()));
  *iters = 1;
  return CQR_OK;
}
size_t cqr_lstsq_workspace_bytes(int64_t m, int64_t, int64_t, int64_t) { return (size_t)m * 100; }
static std::atomic<int> batch_max{0}, batch_calls{0};
static std::atomic<int64_t> batch_wide{0};   // requests with n > 128 seen by a batch
static std::atomic<int64_t> batch_mixed{0};  // batches holding two widths
int cqr_lstsq_batched(cqr_ctx*, int64_t np, const cqr_request* rq, int, double, int* st, int* it, double* up) {
  std::this_thread::sleep_for(std::chrono::milliseconds(solve_ms.load()));
  for (int64_t p = 0; p < np; ++p) {
    st[p] = CQR_OK;
    it[p] = 1;
    up[p] = 0.0;
    if (rq[p].n > 128) ++batch_wide;
    if (rq[p].n != rq[0].n) ++batch_mixed;
  }
  ++batch_calls;
  int cur = batch_max.load();
  while ((int)np > cur && !batch_max.compare_exchange_weak(cur, (int)np)) {}
  return CQR_OK;
}
size_t cqr_lstsq_batched_workspace_bytes(int64_t np, const cqr_request* rq) {
  size_t t = 0;
  for (int64_t p = 0; p < np; ++p) t += (size_t)rq[p].m * 100;
  return t;
}
}

#define CHECK(x)                                         \
  do {                                                   \
    if (!(x)) {                                          \
      printf("FAIL line %d: %s\n", __LINE__, #x);        \
      return 1;                                          \
    }                                                    \
  } while (0)

int main() {
  cqr_server* s = nullptr;
  CHECK(cqr_server_create(&s, 0, 2, 1 << 20) == CQR_OK);
  double d = 0;
  cqr_request r{};
  r.A = &d; r.B = &d; r.X = &d;
  r.m = 10; r.n = 5; r.lda = 10; r.ldb = 10; r.nrhs = 1; r.ldx = 5;
  r.tier = CQR_TIER_F64;
  cqr_result res{};

  // one result, collected once
  const int64_t a = cqr_server_submit(s, &r);
  CHECK(a >= 0);
  CHECK(cqr_server_wait(s, a, &res) == CQR_OK && res.id == a && res.status == CQR_OK);
  CHECK(cqr_server_wait(s, a, &res) == CQR_ERR_INVALID);
  CHECK(cqr_server_wait(s, 999, &res) == CQR_ERR_INVALID);
  CHECK(cqr_server_wait(s, -1, &res) == CQR_ERR_INVALID);

  // drained results cannot be waited for
  const int64_t b = cqr_server_submit(s, &r), c = cqr_server_submit(s, &r);
  cqr_result out[4];
  int total = 0;
  for (int t = 0; t < 1000 && total < 2; ++t) {
    std::this_thread::sleep_for(std::chrono::milliseconds(10));
    int k = 0;
    CHECK(cqr_server_drain(s, out + total, 4 - total, &k) == CQR_OK);
    total += k;
  }
  CHECK(total == 2);
  CHECK(cqr_server_wait(s, b, &res) == CQR_ERR_INVALID);
  CHECK(cqr_server_wait(s, c, &res) == CQR_ERR_INVALID);

  // a waiter racing a drain: exactly one of them collects
  solve_ms = 100;
  const int64_t e = cqr_server_submit(s, &r);
  std::atomic<int> wst{-1};
  std::thread w([&] {
    cqr_result rr;
    wst = cqr_server_wait(s, e, &rr);
  });
  int drained = 0;
  for (int t = 0; t < 1000 && !drained && wst.load() != CQR_OK; ++t) {
    std::this_thread::sleep_for(std::chrono::milliseconds(5));
    int k = 0;
    CHECK(cqr_server_drain(s, out, 4, &k) == CQR_OK);
    drained += k;
  }
  w.join();
  CHECK((wst.load() == CQR_OK) != (drained == 1));
  CHECK(wst.load() == CQR_OK || wst.load() == CQR_ERR_INVALID);

  // two waiters on one id: one result, one CQR_ERR_INVALID
  const int64_t f = cqr_server_submit(s, &r);
  std::atomic<int> s1{-1}, s2{-1};
  std::thread w1([&] {
    cqr_result rr;
    s1 = cqr_server_wait(s, f, &rr);
  });
  std::thread w2([&] {
    cqr_result rr;
    s2 = cqr_server_wait(s, f, &rr);
  });
  w1.join();
  w2.join();
  CHECK((s1.load() == CQR_OK) + (s2.load() == CQR_OK) == 1);
  CHECK(s1.load() == CQR_ERR_INVALID || s2.load() == CQR_ERR_INVALID);

  // a burst through two slots, every result collected once
  solve_ms = 1;
  std::vector<int64_t> ids;
  for (int i = 0; i < 200; ++i) ids.push_back(cqr_server_submit(s, &r));
  for (int64_t id : ids) CHECK(cqr_server_wait(s, id, &res) == CQR_OK && res.id == id && res.status == CQR_OK);
  int64_t queued = -1;
  size_t in_use = 1;
  // the slots' arenas stay resident and are what the server charges
  CHECK(cqr_server_pending(s, &queued, &in_use) == CQR_OK && queued == 0 && in_use == arena_now.load());
  CHECK(in_use > 0 && in_use <= (1u << 20));
  CHECK(cqr_server_destroy(s) == CQR_OK);
  CHECK(arena_now.load() == 0);

  // resident arenas never exceed the budget: two slots, a 3000-byte budget,
  // requests of 1000 and 2000 bytes (arenas 1125 and 2250): an idle slot
  // gives its arena back when the head cannot fit beside it
  {
    arena_peak = 0;
    cqr_server* sb = nullptr;
    CHECK(cqr_server_create(&sb, 0, 2, 3000) == CQR_OK);
    cqr_request small = r, big = r;
    small.m = 10; small.lda = small.ldb = 10;
    big.m = 20; big.lda = big.ldb = 20;
    solve_ms = 2;
    std::vector<int64_t> q;
    for (int i = 0; i < 40; ++i) q.push_back(cqr_server_submit(sb, (i % 3) ? &small : &big));
    for (int64_t id : q) CHECK(cqr_server_wait(sb, id, &res) == CQR_OK && res.status == CQR_OK);
    CHECK(arena_peak.load() <= 3000);
    big.m = 30; big.lda = big.ldb = 30;    // 3000 bytes: its arena (3375) exceeds the budget
    CHECK(cqr_server_submit(sb, &big) == -CQR_ERR_NOMEM);
    CHECK(cqr_server_destroy(sb) == CQR_OK);
  }

  // a thread waiting for a request that is still queued when the server is
  // destroyed returns with CQR_ERR_CANCELLED, and the destroy returns
  {
    cqr_server* sc = nullptr;
    CHECK(cqr_server_create(&sc, 0, 1, 1 << 20) == CQR_OK);
    CHECK(cqr_server_set_batching(sc, 1) == CQR_OK);   // one request per run
    solve_ms = 100;
    const int64_t run = cqr_server_submit(sc, &r), queued_id = cqr_server_submit(sc, &r);
    (void)run;
    std::atomic<int> wst2{-1}, rst{-1};
    std::thread wq([&] {
      cqr_result rr{};
      wst2 = cqr_server_wait(sc, queued_id, &rr);
      rst = rr.status;
    });
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    CHECK(cqr_server_destroy(sc) == CQR_OK);
    wq.join();
    CHECK(wst2.load() == CQR_OK && rst.load() == CQR_ERR_CANCELLED);
  }

  // destroying with requests queued returns once the running one ends
  cqr_server* s1slot = nullptr;
  CHECK(cqr_server_create(&s1slot, 0, 1, 1 << 20) == CQR_OK);
  CHECK(cqr_server_set_batching(s1slot, 1) == CQR_OK);
  solve_ms = 50;
  for (int i = 0; i < 20; ++i) CHECK(cqr_server_submit(s1slot, &r) >= 0);
  const auto t0 = std::chrono::steady_clock::now();
  CHECK(cqr_server_destroy(s1slot) == CQR_OK);
  CHECK(std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(500));
  // batching: one-panel requests queued behind a running one are taken
  // together, at most max_batch at a time and within the budget; a request
  // wider than one panel is never part of a batch
  {
    cqr_server* sd = nullptr;
    CHECK(cqr_server_create(&sd, 0, 1, 1 << 20) == CQR_OK);
    CHECK(cqr_server_set_batching(sd, 8) == CQR_OK);
    CHECK(cqr_server_set_batching(sd, 0) == CQR_ERR_INVALID);
    batch_max = 0;
    batch_calls = 0;
    batch_wide = 0;
    arena_peak = 0;
    solve_ms = 20;
    cqr_request wide = r;
    wide.n = 200; wide.m = 400; wide.lda = wide.ldb = 400; wide.ldx = 200;
    std::vector<int64_t> q;
    q.push_back(cqr_server_submit(sd, &r));
    std::this_thread::sleep_for(std::chrono::milliseconds(5));
    cqr_request other = r;            // one panel too, another width
    other.n = 6; other.ldx = 6;
    for (int i = 0; i < 30; ++i) q.push_back(cqr_server_submit(sd, i == 10 ? &wide : &r));
    for (int i = 0; i < 6; ++i) q.push_back(cqr_server_submit(sd, &other));
    int most = 0, singles_wide = 0;
    for (int64_t id : q) {
      CHECK(cqr_server_wait(sd, id, &res) == CQR_OK && res.status == CQR_OK);
      CHECK(res.batch >= 1 && res.batch <= 8);
      most = res.batch > most ? res.batch : most;
      if (id == q[11]) singles_wide = res.batch;
    }
    CHECK(most == 8 && batch_max.load() == 8);
    CHECK(singles_wide == 1 && batch_wide.load() == 0);
    CHECK(batch_mixed.load() == 0);   // a batch holds one width
    CHECK(batch_calls.load() < 30);
    CHECK(arena_peak.load() <= (1u << 20));
    CHECK(cqr_server_destroy(sd) == CQR_OK);
  }
  // batches stay within the budget: 4000 bytes, one-panel requests of 1000
  // bytes each, so a batch holds at most three (Σ 3000, arena 3375)
  {
    cqr_server* se = nullptr;
    CHECK(cqr_server_create(&se, 0, 1, 4000) == CQR_OK);
    batch_max = 0;
    arena_peak = 0;
    solve_ms = 5;
    std::vector<int64_t> q;
    for (int i = 0; i < 20; ++i) q.push_back(cqr_server_submit(se, &r));
    for (int64_t id : q) CHECK(cqr_server_wait(se, id, &res) == CQR_OK && res.status == CQR_OK);
    CHECK(batch_max.load() <= 3 && arena_peak.load() <= 4000);
    CHECK(cqr_server_destroy(se) == CQR_OK);
  }
  printf("server logic ok\n");
  return 0;
}
