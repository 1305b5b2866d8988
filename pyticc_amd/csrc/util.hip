// Error reporting, workspace management and small layout kernels.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "ticc_common.h"

namespace ticc {

namespace {
thread_local char g_error[1024] = "";

struct Buffer {
  void* ptr = nullptr;
  size_t size = 0;
};
constexpr int kMaxDevices = 64;
// Library state is kept per (device, lane): a host thread that selects its own
// lane (set_lane) gets workspaces, auxiliary streams and events no other lane
// touches, so sweeps on different lanes may run concurrently.
Buffer g_ws[kMaxDevices][TICC_LANES][WS_COUNT];
std::mutex g_ws_mutex;
constexpr int kMaxAux = 16;
hipStream_t g_streams[kMaxDevices][TICC_LANES][kMaxAux] = {};
hipEvent_t g_events[kMaxDevices][TICC_LANES][kMaxAux * 4] = {};
thread_local int t_lane = 0;

__global__ void k_pad_identity(const double* __restrict__ src, long long src_stride, int N, double* __restrict__ dst,
                               int Np) {
  const int b = blockIdx.y;
  const size_t n2 = (size_t)Np * Np;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n2; idx += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(idx / Np);
    const int c = (int)(idx % Np);
    double v;
    if (r < N && c < N)
      v = src[b * src_stride + (size_t)r * N + c];
    else
      v = (r == c) ? 1.0 : 0.0;
    dst[b * n2 + idx] = v;
  }
}

__global__ void k_unpad(const double* __restrict__ src, int Np, double* __restrict__ dst, long long dst_stride, int N) {
  const int b = blockIdx.y;
  const size_t n2 = (size_t)N * N;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n2; idx += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(idx / N);
    const int c = (int)(idx % N);
    dst[b * dst_stride + idx] = src[(size_t)b * Np * Np + (size_t)r * Np + c];
  }
}

__global__ void k_check_finite(const double* __restrict__ x, long long n, int* __restrict__ flag) {
  int bad = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    if (!isfinite(x[i])) bad = 1;
  if (bad) atomicOr(flag, 1);
}

int grid_for(long long n) {
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  return (int)blocks;
}
}  // namespace

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
  return code;
}

const char* last_error() { return g_error; }

int set_lane(int lane) {
  if (lane < 0 || lane >= TICC_LANES) return fail(TICC_ERR_VALUE, "lane must lie in [0, %d), got %d", TICC_LANES, lane);
  t_lane = lane;
  return TICC_OK;
}

int workspace(WorkspaceSlot slot, size_t bytes, void** out) {
  int dev = 0;
  TICC_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(TICC_ERR_VALUE, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  Buffer& buf = g_ws[dev][t_lane][slot];
  if (buf.size < bytes) {
    if (buf.ptr) {
      TICC_CHECK_HIP(hipDeviceSynchronize());
      TICC_CHECK_HIP(hipFree(buf.ptr));
      buf.ptr = nullptr;
      buf.size = 0;
    }
    size_t want = bytes + bytes / 8;
    hipError_t e = hipMalloc(&buf.ptr, want);
    if (e != hipSuccess) {
      buf.ptr = nullptr;
      (void)hipGetLastError();
      return fail(TICC_ERR_NOMEM, "workspace allocation of %zu bytes failed: %s", want, hipGetErrorString(e));
    }
    buf.size = want;
  }
  *out = buf.ptr;
  return TICC_OK;
}

int aux_streams(int n, hipStream_t* out) {
  int dev = 0;
  TICC_CHECK_HIP(hipGetDevice(&dev));
  if (n > kMaxAux || dev >= kMaxDevices) return fail(TICC_ERR_VALUE, "too many auxiliary streams requested");
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  for (int i = 0; i < n; ++i) {
    hipStream_t& st = g_streams[dev][t_lane][i];
    if (!st) TICC_CHECK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    out[i] = st;
  }
  return TICC_OK;
}

int aux_events(int n, hipEvent_t* out) {
  int dev = 0;
  TICC_CHECK_HIP(hipGetDevice(&dev));
  if (n > kMaxAux * 4 || dev >= kMaxDevices) return fail(TICC_ERR_VALUE, "too many auxiliary events requested");
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  for (int i = 0; i < n; ++i) {
    hipEvent_t& ev = g_events[dev][t_lane][i];
    if (!ev) TICC_CHECK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    out[i] = ev;
  }
  return TICC_OK;
}

int release_workspace() {
  int dev = 0;
  TICC_CHECK_HIP(hipGetDevice(&dev));
  TICC_CHECK_HIP(hipDeviceSynchronize());
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  for (int l = 0; l < TICC_LANES; ++l)
    for (int s = 0; s < WS_COUNT; ++s) {
      Buffer& buf = g_ws[dev][l][s];
      if (buf.ptr) TICC_CHECK_HIP(hipFree(buf.ptr));
      buf.ptr = nullptr;
      buf.size = 0;
    }
  return TICC_OK;
}

int pad_identity(const double* src, long long src_stride, int N, double* dst, int Np, int batch, hipStream_t stream) {
  hipLaunchKernelGGL(k_pad_identity, dim3(grid_for((long long)Np * Np), batch), dim3(256), 0, stream, src,
                     src_stride, N, dst, Np);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int unpad(const double* src, int Np, double* dst, long long dst_stride, int N, int batch, hipStream_t stream) {
  hipLaunchKernelGGL(k_unpad, dim3(grid_for((long long)N * N), batch), dim3(256), 0, stream, src, Np, dst, dst_stride,
                     N);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int check_finite(const double* x, long long n, int* flag, hipStream_t stream) {
  hipLaunchKernelGGL(k_check_finite, dim3(grid_for(n)), dim3(256), 0, stream, x, n, flag);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

}  // namespace ticc
