// TSQR + Householder reconstruction of one panel without any allocation:
// cqr_hh_recon_qr_noalloc runs the kernels of cqr_hh_recon_qr_ex on a
// fixed scratch buffer, a file-scope __device__ array, so a caller that
// cannot allocate (code under HIP graph capture, a serving loop that must
// not call hipMalloc) can factor panels whose workspace fits it.
//
// Each GPU holds its own copy of the array from module load to process
// end.  Its address on a device is looked up once, with that device
// current, when a context is created there (cqr_create calls
// noalloc_setup); host code uses that address, never the array's name.
// Calls on one GPU share the array: calls that overlap in time race, and
// every call overwrites what the previous one left (nothing in it needs to
// start at zero: each kernel writes its outputs before they are read).
#include <mutex>

#include "cqr_internal.h"

namespace cqr {

namespace {

constexpr size_t kScratchBytes = size_t(96) << 20;   // 96 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

size_t noalloc_scratch_bytes() { return kScratchBytes; }

// Device address of this GPU's scratch array; called with c->device current.
int noalloc_setup(cqr_ctx* c) {
  if (c->device < 0 || c->device >= kMaxDevices) return set_err(c, CQR_ERR_DEVICE, "device %d out of range", c->device);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[c->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return set_err(c, CQR_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    }
    g_ptr[c->device] = p;
  }
  c->noalloc = g_ptr[c->device];
  return CQR_OK;
}

}  // namespace cqr

using namespace cqr;

extern "C" {

size_t cqr_noalloc_scratch_bytes(void) { return kScratchBytes; }

int cqr_hh_recon_qr_noalloc(cqr_ctx* c, double* A, int64_t lda, int64_t m, int64_t b, double* W, double* T,
                            double* invT_T, int64_t nleaves) {
  CQR_CHECK_CTX(c);
  if (!A || !W || b < 1 || m < b || lda < m || nleaves < 0)
    return set_err(c, CQR_ERR_INVALID, "hh_recon_qr_noalloc: bad arguments (m=%lld b=%lld lda=%lld)",
                   (long long)m, (long long)b, (long long)lda);
  if (b > CQR_MAX_B) return set_err(c, CQR_ERR_UNSUPPORTED, "hh_recon_qr_noalloc: b=%lld > %d", (long long)b, CQR_MAX_B);
  const size_t need = panel_ws_bytes(m, b, nleaves, c->leaf_rows);
  if (!need) return set_err(c, CQR_ERR_UNSUPPORTED, "hh_recon_qr_noalloc: no leaf plan");
  if (need > kScratchBytes)
    return set_err(c, CQR_ERR_UNSUPPORTED,
                   "hh_recon_qr_noalloc: the %lld×%lld panel needs %zu bytes of scratch, more than the fixed %zu",
                   (long long)m, (long long)b, need, kScratchBytes);
  if (!c->noalloc) return set_err(c, CQR_ERR_DEVICE, "hh_recon_qr_noalloc: no scratch address on this device");
  return hh_recon_panel(c, A, lda, m, b, W, T, invT_T, nullptr, nleaves, static_cast<char*>(c->noalloc), nullptr);
}

}  // extern "C"
