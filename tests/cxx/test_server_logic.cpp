// Host-only test of the request server's queue and collection logic
// (candmc_amd/csrc/server.cpp): the GPU calls it makes are replaced by the
// stand-ins below, so each "solve" is a short sleep.  Checked: a result is
// collected exactly once; a wait on an id never submitted, or already
// collected by a wait or a drain, returns CQR_ERR_INVALID instead of
// blocking; two waiters on one id, or a waiter racing a drain, end with
// exactly one collector; every request of a burst through two slots is
// served; destroying a server with requests queued drops them and returns
// once the running one ends.  Built with g++ and run by tests/test_server_logic.py.
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
int cqr_create(cqr_ctx** c, int) {
  *c = static_cast<cqr_ctx*>(malloc(8));
  return CQR_OK;
}
int cqr_destroy(cqr_ctx* c) {
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
size_t cqr_lstsq_workspace_bytes(int64_t, int64_t, int64_t, int64_t) { return 1000; }
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
  CHECK(cqr_server_pending(s, &queued, &in_use) == CQR_OK && queued == 0 && in_use == 0);
  CHECK(cqr_server_destroy(s) == CQR_OK);

  // destroying with requests queued returns once the running one ends
  cqr_server* s1slot = nullptr;
  CHECK(cqr_server_create(&s1slot, 0, 1, 1 << 20) == CQR_OK);
  solve_ms = 50;
  for (int i = 0; i < 20; ++i) CHECK(cqr_server_submit(s1slot, &r) >= 0);
  const auto t0 = std::chrono::steady_clock::now();
  CHECK(cqr_server_destroy(s1slot) == CQR_OK);
  CHECK(std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(500));
  printf("server logic ok\n");
  return 0;
}
