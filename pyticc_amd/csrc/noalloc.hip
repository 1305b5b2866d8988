// Batched inversion that allocates nothing: its scratch is a file-scope device
// array of fixed size, one copy per GPU, held until the process ends.
//
// For callers that cannot call hipMalloc (code under stream capture, tight
// serving loops, harnesses that forbid allocation).  The array's host-side name
// is only a stand-in for the symbol: host code uses the device address that
// hipGetSymbolAddress returns for the current device, fetched once per device
// and kept.  Every call on one GPU shares that GPU's array, so calls that
// overlap on one GPU (different streams or host threads) race; the kernels
// write every scratch element they read, so nothing needs clearing between
// calls.
#include <mutex>

#include "ticc_common.h"

namespace {

__device__ __align__(256) char scratch[TICC_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (looked up once per device).
int scratch_address(void** out) {
  int dev = 0;
  TICC_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return ticc::fail(TICC_ERR_UNSUPPORTED, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    TICC_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return ticc::fail(TICC_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TICC_OK;
}

}  // namespace

extern "C" {

int ticc_invert_batched_noalloc(int n_batch, int n_channel, double* A, void* stream) {
  if (n_channel < 1 || n_channel > TICC_MAX_CHANNELS)
    return ticc::fail(TICC_ERR_VALUE, "invert_batched_noalloc: n_channel=%d outside [1, %d]", n_channel,
                      TICC_MAX_CHANNELS);
  if (n_batch < 1) return ticc::fail(TICC_ERR_VALUE, "invert_batched_noalloc: n_batch=%d must be positive", n_batch);
  const size_t need = ticc::invert_scratch_bytes(n_batch, ticc::round_up(n_channel, ticc::GJ_NB));
  if (need > (size_t)TICC_NOALLOC_SCRATCH_BYTES)
    return ticc::fail(TICC_ERR_VALUE,
                      "invert_batched_noalloc: %d matrices of %d channels need %zu bytes of scratch, above the fixed "
                      "%d bytes (use ticc_invert_batched, or fewer matrices per call)",
                      n_batch, n_channel, need, (int)TICC_NOALLOC_SCRATCH_BYTES);
  if (!A) return ticc::fail(TICC_ERR_VALUE, "invert_batched_noalloc: null matrix pointer");
  void* base = nullptr;
  TICC_RETURN_IF(scratch_address(&base));
  return ticc::invert_with_scratch(n_batch, n_channel, A, base, static_cast<hipStream_t>(stream));
}

}  // extern "C"
