// Batched FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
// C[b] = alpha op(A[b]) op(B[b]) + beta C[b] for padded row-major matrices:
// m, n multiples of 32, k a multiple of 4.  A workgroup of 4 waves owns a
// 64 x 64 tile of C, each wave a 32 x 32 sub-tile (2 x 2 MFMA blocks).
// Operands are read straight from global memory (L2-resident at the sizes the
// matching stage uses); this GEMM is off the LDMD sweep's critical path.
#include "ticc_common.h"

namespace ticc {
namespace {

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_dgemm(int m, int n, int k, double alpha, const double* __restrict__ A,
                                               int lda, long long sA, const double* __restrict__ B, int ldb,
                                               long long sB, double beta, double* __restrict__ C, int ldc,
                                               long long sC) {
  const int tn = (n + 63) / 64;
  const int i0 = (blockIdx.x / tn) * 64 + ((threadIdx.x >> 6) >> 1) * 32;
  const int j0 = (blockIdx.x % tn) * 64 + ((threadIdx.x >> 6) & 1) * 32;
  if (i0 >= m || j0 >= n) return;
  const int b = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int lr = lane & 15;
  const int lq = lane >> 4;
  const double* Ab = A + b * sA;
  const double* Bb = B + b * sB;
  double* Cb = C + b * sC;

  double4_t acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = double4_t{0.0, 0.0, 0.0, 0.0};

  for (int kk = 0; kk < k; kk += 4) {
    const int kq = kk + lq;
    double a[2], bv[2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
      const int row = i0 + 16 * bi + lr;
      a[bi] = TA ? Ab[(size_t)kq * lda + row] : Ab[(size_t)row * lda + kq];
    }
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      const int col = j0 + 16 * bj + lr;
      bv[bj] = TB ? Bb[(size_t)col * ldb + kq] : Bb[(size_t)kq * ldb + col];
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
        acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], bv[bj], acc[bi][bj], 0, 0, 0);
  }

#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + 16 * bi + mfma_drow(lq, r);
        const int col = j0 + 16 * bj + lr;
        double* cp = Cb + (size_t)row * ldc + col;
        const double v = alpha * acc[bi][bj][r];
        *cp = (beta == 0.0) ? v : v + beta * (*cp);
      }
}

}  // namespace

int dgemm_batched(bool transA, bool transB, int m, int n, int k, double alpha, const double* A, int lda,
                  long long strideA, const double* B, int ldb, long long strideB, double beta, double* C, int ldc,
                  long long strideC, int batch, hipStream_t stream) {
  if (m <= 0 || n <= 0 || k <= 0 || batch <= 0 || m % 32 || n % 32 || k % 4)
    return fail(TICC_ERR_VALUE, "dgemm_batched: unsupported shape m=%d n=%d k=%d batch=%d", m, n, k, batch);
  dim3 grid(((m + 63) / 64) * ((n + 63) / 64), batch);
  if (!transA && !transB)
    hipLaunchKernelGGL((k_dgemm<false, false>), grid, dim3(256), 0, stream, m, n, k, alpha, A, lda, strideA, B, ldb,
                       strideB, beta, C, ldc, strideC);
  else if (transA && !transB)
    hipLaunchKernelGGL((k_dgemm<true, false>), grid, dim3(256), 0, stream, m, n, k, alpha, A, lda, strideA, B, ldb,
                       strideB, beta, C, ldc, strideC);
  else if (!transA && transB)
    hipLaunchKernelGGL((k_dgemm<false, true>), grid, dim3(256), 0, stream, m, n, k, alpha, A, lda, strideA, B, ldb,
                       strideB, beta, C, ldc, strideC);
  else
    hipLaunchKernelGGL((k_dgemm<true, true>), grid, dim3(256), 0, stream, m, n, k, alpha, A, lda, strideA, B, ldb,
                       strideB, beta, C, ldc, strideC);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

}  // namespace ticc
