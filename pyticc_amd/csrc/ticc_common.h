// Shared definitions for the CDNA4 (gfx950) coupled-channel kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/ticc_hip.h"

namespace ticc {

// Record a formatted error message for ticc_last_error(); returns `code`.
int fail(int code, const char* fmt, ...);

// Device-memory workspace owned by the library, one growable buffer per
// (device, slot).  Slots let the contraction and the sweep run concurrently
// from two host threads without sharing scratch memory.
enum WorkspaceSlot { WS_SWEEP = 0, WS_CONTRACT = 1, WS_MATCH = 2, WS_INVERT = 3, WS_COUNT = 4 };
int workspace(WorkspaceSlot slot, size_t bytes, void** out);
// Select the calling thread's lane of library state (workspaces, auxiliary
// streams and events), 0 <= lane < TICC_LANES; lane 0 by default.
int set_lane(int lane);
// Library-owned non-blocking streams (default priority) and events of the
// current device, created once and reused; callers must not destroy them.
int aux_streams(int n, hipStream_t* out);
int aux_events(int n, hipEvent_t* out);

// Panel width of the blocked Gauss-Jordan inversion and the padding quantum of
// every internal square matrix (Np = round_up(N, GJ_NB)).
constexpr int GJ_NB = 32;
constexpr int MAX_NP = 1536;  // three panel rows per thread of a 512-thread workgroup
// Scratch of gj_invert_raw per matrix: pout holds 9 Np x NB panels (two
// transformed panels, the look-ahead columns of one and of two steps, and the
// 2 NB transposed update coefficients of two pairs), src 4 Np row maps, piv 6
// NB-entry row lists.
constexpr int GJ_POUT_PANELS = 9;
constexpr int GJ_SRC_MAPS = 4;
constexpr int GJ_PIV_LISTS = 6;

inline int round_up(int n, int q) { return (n + q - 1) / q * q; }

}  // namespace ticc

#define TICC_CHECK_HIP(call)                                                                  \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return ticc::fail(TICC_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                  \
  } while (0)

#define TICC_CHECK_LAUNCH() TICC_CHECK_HIP(hipGetLastError())

#define TICC_RETURN_IF(...)             \
  do {                                  \
    int s_ = (__VA_ARGS__);             \
    if (s_ != TICC_OK) return s_;       \
  } while (0)

typedef double double4_t __attribute__((ext_vector_type(4)));

// Output layout of v_mfma_f64_16x16x4f64 on gfx950: lane l holds A[l%16][l/16]
// and B[l/16][l%16] as inputs; accumulator register r of lane l is
// D[mfma_drow(l/16, r)][l%16].
__host__ __device__ constexpr int mfma_drow(int lq, int r) { return lq + 4 * r; }

typedef float float4_t __attribute__((ext_vector_type(4)));

namespace ticc {
// Element-type helpers of the kernels that run in FP64 or FP32 (the
// Gauss-Jordan inversion, the refinement GEMMs): the 16x16x4 MFMA of the type,
// whose output row of accumulator register r in lane group lq differs between
// the two, FMA and two-element vectors.
template <typename T>
struct Elem;
template <>
struct Elem<double> {
  typedef double4_t acc;
  typedef double2 vec2;
  __device__ static __forceinline__ acc mfma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ constexpr int drow(int lq, int r) { return mfma_drow(lq, r); }
  __device__ static __forceinline__ double fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
  __device__ static __forceinline__ vec2 pair(double x, double y) { return make_double2(x, y); }
};
template <>
struct Elem<float> {
  typedef float4_t acc;
  typedef float2 vec2;
  __device__ static __forceinline__ acc mfma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // v_mfma_f32_16x16x4f32: lane l, register r holds D[4 (l / 16) + r][l % 16]
  __device__ static __forceinline__ constexpr int drow(int lq, int r) { return 4 * lq + r; }
  __device__ static __forceinline__ float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
  __device__ static __forceinline__ vec2 pair(float x, float y) { return make_float2(x, y); }
};

}  // namespace ticc

// Internal entry points shared between translation units.
namespace ticc {
// Blocked GJ inversion of `batch` padded Np x Np matrices in place in bufA
// (bufB is the ping-pong partner).  On return the inverse, up to a pending
// column permutation, is in *result (== bufA or bufB) and pinv[b*Np + c] gives
// the column of *result that holds column c of the inverse.
int gj_invert_raw(double* bufA, double* bufB, int batch, int Np, double** result, int* perm, int* pinv,
                  double* pout, int* src, int* piv, hipStream_t stream);
// The same inversion restricted to the matrices b with mask[b] != 0 (device
// array): every kernel's workgroups of the other matrices exit at once.
int gj_invert_raw_masked(double* bufA, double* bufB, int batch, int Np, double** result, int* perm, int* pinv,
                         double* pout, int* src, int* piv, const int* mask, hipStream_t stream);
// The same inversion in FP32 (FP32 panels, FP32 MFMA updates, same pivot rule).
int gj_invert_raw_f32(float* bufA, float* bufB, int batch, int Np, float** result, int* perm, int* pinv, float* pout,
                      int* src, int* piv, hipStream_t stream);
// Gather the column permutation into a plain inverse: out[b][r][c] = in[b][r][pinv[c]].
int gj_unpermute(const double* in, double* out, int batch, int Np, const int* pinv, hipStream_t stream);
// The gather for the matrices with mask[b] != 0 only (the others keep `out`).
int gj_unpermute_masked(const double* in, double* out, int batch, int Np, const int* pinv, const int* mask,
                        hipStream_t stream);
// The same gather from an FP32 inverse into FP64.
int gj_unpermute_f32(const float* in, double* out, int batch, int Np, const int* pinv, hipStream_t stream);
int gj_unpermute_f32f32(const float* in, float* out, int batch, int Np, const int* pinv, hipStream_t stream);
// Batched C = alpha op(A) op(B) + beta C for padded matrices (m, n, k multiples of 32).
int dgemm_batched(bool transA, bool transB, int m, int n, int k, double alpha, const double* A, int lda,
                  long long strideA, const double* B, int ldb, long long strideB, double beta, double* C, int ldc,
                  long long strideC, int batch, hipStream_t stream);
// Batched square GEMMs of the Newton refinement (Np a multiple of 32):
// out = beta Cin + alpha A B + diag I, out32 = (float) out; either output may be
// null.  FP64 operands on the FP64 MFMA, or FP32 operands on the FP32 MFMA.  The
// FP64 form can raise rho[b] (float bits) to the largest |out[b]| element.
int refine_gemm_f64(int Np, int batch, const double* A, const double* B, double alpha, const double* Cin, double beta,
                    double diag, double* out, float* out32, unsigned* rho, hipStream_t stream);
int refine_gemm_f32(int Np, int batch, const float* A, const float* B, double alpha, const double* Cin, double beta,
                    double diag, double* out, float* out32, hipStream_t stream);
// Enqueue-only form of the LDMD sweep (for stream capture): device-resident
// energies (one per batch item) and half steps, the initial Y read from Y_in,
// caller-owned scratch (scratch == null: only report the bytes needed), and no
// host copy, memset, synchronisation or finite check.
struct SweepAsync {
  const double* E_dev;
  const double* hs_dev;
  const double* Y_in;
  void* scratch;
  size_t scratch_bytes;
};
int ldmd_propagate(int nE, int N, const double* E_host, double mu, int nS, const double* hs_host, const double* Ws0,
                   const double* Wm0, const double* We0, long long wstride, double* Y, hipStream_t stream, int half,
                   int n_block, long long gstride, int refine, SweepAsync* as);
// FP32 copy of n doubles.
int to_f32(const double* in, float* out, long long n, hipStream_t stream);
// Identity padding: dst (batch, Np, Np) <- src (batch, N, N) with ld N, identity in the padding.
int pad_identity(const double* src, long long src_stride, int N, double* dst, int Np, int batch, hipStream_t stream);
int unpad(const double* src, int Np, double* dst, long long dst_stride, int N, int batch, hipStream_t stream);
// Batched inversion of n_batch N x N matrices in place (pad, Gauss-Jordan, column
// unpermute, unpad), its scratch carved from `scratch` (invert_scratch_bytes of
// it); synchronises `st`.  Shared by the workspace and the no-allocation entries.
size_t invert_scratch_bytes(int n_batch, int Np);
int invert_with_scratch(int n_batch, int n_channel, double* A, void* scratch, hipStream_t st);
// Non-finite scan: *flag (device int) set to 1 if any of n doubles is NaN/Inf.
int check_finite(const double* x, long long n, int* flag, hipStream_t stream);
}  // namespace ticc
