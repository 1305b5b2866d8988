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

// Batched square GEMM of the Newton refinement of an approximate inverse
// (Np x Np row-major, Np a multiple of 32):
//   out[b] = beta Cin[b] + alpha A[b] B[b] + diag I,   out32[b] = (float) out[b],
// either output optional; with rho set, rho[b] becomes at least the largest
// |out[b]| element (as the bits of a non-negative float, which order as unsigned
// integers).  TI is the operand type: FP64 operands on the FP64
// MFMA, FP32 operands on the FP32 MFMA with FP32 accumulation.  A workgroup of
// four waves owns a 64 x 64 tile of out (32 x 32 per wave) and walks K in slabs
// of 32 staged through LDS, double-buffered: the next slab's global loads are
// in flight while the current slab's MFMAs run.  With a batch that is a
// multiple of 8 the tiles of one matrix share an XCD (workgroups are dealt to
// the 8 XCDs round robin), so its operand rows are re-read from that XCD's L2.
template <typename TI>
__global__ __launch_bounds__(256) void k_gemm_sq(int Np, int tiles, const TI* __restrict__ A, const TI* __restrict__ B,
                                                 double alpha, const double* __restrict__ Cin, double beta, double diag,
                                                 double* __restrict__ out, float* __restrict__ out32,
                                                 unsigned* __restrict__ rho) {
  typedef Elem<TI> E;
  constexpr int KS = 32;
  __shared__ TI As[2][64][KS + 1];
  __shared__ TI Bs[2][KS][64 + 4];
  const int per = tiles * tiles;
  const int lin = blockIdx.x + gridDim.x * blockIdx.y;
  int b, tile;
  if ((gridDim.y & 7) == 0) {
    const int k = lin >> 3;
    b = (lin & 7) + 8 * (k / per);
    tile = k % per;
  } else {
    b = blockIdx.y;
    tile = blockIdx.x;
  }
  const int r0 = (tile / tiles) * 64, c0 = (tile % tiles) * 64;
  const int t = threadIdx.x;
  const int wave = t >> 6, lane = t & 63, lr = lane & 15, lq = lane >> 4;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const size_t n2 = (size_t)Np * Np;
  const TI* Ab = A + b * n2;
  const TI* Bb = B + b * n2;
  // global -> register staging: A rows (t / 4), columns 8 (t % 4) .. + 8 of the
  // slab; B slab row t / 8, columns 8 (t % 8) .. + 8 of the tile
  const int ar = t >> 2, ac = (t & 3) * 8;
  const int br = t >> 3, bc = (t & 7) * 8;
  const bool a_ok = r0 + ar < Np;
  const bool b_ok = c0 + bc < Np;  // Np and c0 are multiples of 32: 8 columns are all in or all out
  TI ra[8], rb[8];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) ra[i] = a_ok ? Ab[(size_t)(r0 + ar) * Np + k0 + ac + i] : TI(0);
#pragma unroll
    for (int i = 0; i < 8; ++i) rb[i] = b_ok ? Bb[(size_t)(k0 + br) * Np + c0 + bc + i] : TI(0);
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int i = 0; i < 8; ++i) As[buf][ar][ac + i] = ra[i];
#pragma unroll
    for (int i = 0; i < 8; ++i) Bs[buf][br][bc + i] = rb[i];
  };
  typename E::acc acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[bi][bj][r] = 0;
  const int nk = Np / KS;
  load(0);
  stage(0);
  __syncthreads();
  for (int ks = 0; ks < nk; ++ks) {
    const int buf = ks & 1;
    if (ks + 1 < nk) load((ks + 1) * KS);
#pragma unroll
    for (int kk = 0; kk < KS / 4; ++kk) {
      const int kq = 4 * kk + lq;
      TI av[2], bv[2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) av[bi] = As[buf][wr + 16 * bi + lr][kq];
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) bv[bj] = Bs[buf][kq][wc + 16 * bj + lr];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = E::mfma(av[bi], bv[bj], acc[bi][bj]);
    }
    if (ks + 1 < nk) stage(buf ^ 1);
    __syncthreads();
  }
  float vmax = 0.0f;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = r0 + wr + 16 * bi + E::drow(lq, r);
      if (i >= Np) continue;
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) {
        const int j = c0 + wc + 16 * bj + lr;
        if (j >= Np) continue;
        const size_t o = b * n2 + (size_t)i * Np + j;
        double v = alpha * (double)acc[bi][bj][r];
        if (beta != 0.0) v += beta * Cin[o];
        if (i == j) v += diag;
        if (out) out[o] = v;
        if (out32) out32[o] = (float)v;
        vmax = fmaxf(vmax, (float)fabs(v));
      }
    }
  if (rho) {  // one atomic per wave
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
    if (lane == 0) atomicMax(rho + b, __float_as_uint(vmax));
  }
}

template <typename TI>
int gemm_sq(int Np, int batch, const TI* A, const TI* B, double alpha, const double* Cin, double beta, double diag,
            double* out, float* out32, unsigned* rho, hipStream_t stream) {
  if (Np <= 0 || Np % 32 || batch <= 0 || (beta != 0.0 && !Cin) || (!out && !out32))
    return fail(TICC_ERR_VALUE, "gemm_sq: unsupported arguments Np=%d batch=%d", Np, batch);
  const int tiles = (Np + 63) / 64;
  hipLaunchKernelGGL((k_gemm_sq<TI>), dim3(tiles * tiles, batch), dim3(256), 0, stream, Np, tiles, A, B, alpha, Cin,
                     beta, diag, out, out32, rho);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

}  // namespace

int refine_gemm_f64(int Np, int batch, const double* A, const double* B, double alpha, const double* Cin, double beta,
                    double diag, double* out, float* out32, unsigned* rho, hipStream_t stream) {
  return gemm_sq<double>(Np, batch, A, B, alpha, Cin, beta, diag, out, out32, rho, stream);
}

int refine_gemm_f32(int Np, int batch, const float* A, const float* B, double alpha, const double* Cin, double beta,
                    double diag, double* out, float* out32, hipStream_t stream) {
  return gemm_sq<float>(Np, batch, A, B, alpha, Cin, beta, diag, out, out32, nullptr, stream);
}

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
