// BF16 multilayer perceptron kernels of the learned PES surrogate (gfx950).
//
// The network maps 5 features of a geometry to the interaction energy:
// 5 -> 256 -> 256 -> 256 -> 256 -> 1 with SiLU after each hidden layer.  Weights
// and activations are BF16 MFMA operands (v_mfma_f32_16x16x16bf16_1k), every
// product accumulates in FP32, and the FP32 master weights are updated by Adam.
//
// Shapes are padded to the GEMM tile: the 5 features to 64 columns (zeros), the
// one output to 64 columns.  A layer is C = A . B^T with both operands row-major
// and K-contiguous, which covers the three products of training:
//   forward   Z_l  = A_{l-1} . W_l^T        (A: n x din,  B = W_l:    dout x din)
//   backward  dA   = dZ_l . W_l = dZ_l . (W_l^T)^T
//                                           (A: n x dout, B = W_l^T:  din x dout)
//   gradient  dW_l = dZ_l^T . A_{l-1}       (A = dZ_l^T:  dout x n,
//                                            B = A_{l-1}^T: din x n)
// so activations and output gradients are written both row-major and transposed.
#include <cstdint>

#include "ticc_common.h"

namespace ticc {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even; NaN stays NaN.
__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

enum { EPI_STORE = 0, EPI_HIDDEN = 1, EPI_OUTPUT = 2, EPI_BACK = 3 };

// C[M][N] = A[M][K] . B[N][K]^T.  A workgroup of 4 waves owns a 64 x 64 tile of
// C, each wave a 32 x 32 sub-tile (2 x 2 MFMA blocks of 16 x 16); lane l holds
// A[m = l % 16][k = 4 (l / 16) + i] and B^T[k][n = l % 16], and its results are
// C[4 (l / 16) + i][l % 16].  Epilogues:
//   EPI_STORE   Cf = C                                  (weight gradients)
//   EPI_HIDDEN  Z = C + bias -> Cf; a = SiLU(Z) -> Cb (bf16) and CbT (bf16, N x M)
//   EPI_OUTPUT  Cf = C + bias                           (network output)
//   EPI_BACK    dZ = C * SiLU'(Zin) -> Cf, Cb and CbT    (back through a hidden layer)
//
// Split form (SPLIT): each operand is the sum of a BF16 high part and a BF16 low
// part (x_lo = bf16(x - x_hi)), and C = A_hi B_hi + A_hi B_lo + A_lo B_hi (three
// MFMAs per step, FP32 accumulation): products to about 16 significant bits
// instead of 8.  A_lo may be null (an input without a low part).  EPI_HIDDEN then
// also writes the activation's low part Cb_lo.

template <int EPI, bool SPLIT = false>
__global__ __launch_bounds__(256) void k_mlp_gemm(int M, int N, int K, const uint16_t* __restrict__ A,
                                                  const uint16_t* __restrict__ B, const float* __restrict__ bias,
                                                  const float* __restrict__ Zin, float* __restrict__ Cf,
                                                  uint16_t* __restrict__ Cb, uint16_t* __restrict__ CbT,
                                                  const uint16_t* __restrict__ A_lo = nullptr,
                                                  const uint16_t* __restrict__ B_lo = nullptr,
                                                  uint16_t* __restrict__ Cb_lo = nullptr) {
  const int tn = N / 64;
  const int m0 = (blockIdx.x / tn) * 64, n0 = (blockIdx.x % tn) * 64;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  const int wm = m0 + (wave >> 1) * 32, wn = n0 + (wave & 1) * 32;
  floatx4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a[2], b[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      a[t] = *reinterpret_cast<const bf16x4*>(A + (size_t)(wm + 16 * t + r) * K + k0 + 4 * q);
      b[t] = *reinterpret_cast<const bf16x4*>(B + (size_t)(wn + 16 * t + r) * K + k0 + 4 * q);
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[bi], b[bj], acc[bi][bj], 0, 0, 0);
    if constexpr (SPLIT) {
      bf16x4 al[2], bl[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        bl[t] = *reinterpret_cast<const bf16x4*>(B_lo + (size_t)(wn + 16 * t + r) * K + k0 + 4 * q);
        if (A_lo) al[t] = *reinterpret_cast<const bf16x4*>(A_lo + (size_t)(wm + 16 * t + r) * K + k0 + 4 * q);
      }
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[bi], bl[bj], acc[bi][bj], 0, 0, 0);
          if (A_lo) acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[bi], b[bj], acc[bi][bj], 0, 0, 0);
        }
    }
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = wm + 16 * bi + 4 * q + i;
        const int col = wn + 16 * bj + r;
        const size_t o = (size_t)row * N + col;
        const float v = acc[bi][bj][i];
        if (EPI == EPI_STORE) {
          Cf[o] = v;
        } else if (EPI == EPI_OUTPUT) {
          Cf[o] = v + bias[col];
        } else if (EPI == EPI_HIDDEN) {
          const float z = v + bias[col];
          Cf[o] = z;
          const float act = z * sigmoid(z);
          const uint16_t h = to_bf16(act);
          Cb[o] = h;
          CbT[(size_t)col * M + row] = h;
          if constexpr (SPLIT) Cb_lo[o] = to_bf16(act - __uint_as_float((uint32_t)h << 16));
        } else {
          const float z = Zin[o];
          const float s = sigmoid(z);
          const float dz = v * s * (1.0f + z * (1.0f - s));
          Cf[o] = dz;
          const uint16_t h = to_bf16(dz);
          Cb[o] = h;
          CbT[(size_t)col * M + row] = h;
        }
      }
}

// Features of sample i, (R, theta_X, theta_Y, phi) -> [r_scale / R, cos theta_X,
// cos theta_Y, cos phi, sin theta_X sin theta_Y cos phi], written in BF16 to
// row i of an n x 64 matrix (columns 5..63 zero) and to its transpose.
__device__ __forceinline__ void put_features(int i, int n, double R, double tx, double ty, double ph, double r_scale,
                                             uint16_t* __restrict__ F, uint16_t* __restrict__ FT) {
  const double cph = cos(ph);
  const float f[5] = {(float)(r_scale / R), (float)cos(tx), (float)cos(ty), (float)cph,
                      (float)(sin(tx) * sin(ty) * cph)};
#pragma unroll
  for (int c = 0; c < 64; ++c) {
    const uint16_t h = c < 5 ? to_bf16(f[c]) : (uint16_t)0;
    F[(size_t)i * 64 + c] = h;
    if (FT) FT[(size_t)c * n + i] = h;
  }
}

__global__ void k_mlp_features(int n, int n_valid, const double* __restrict__ X4, double r_scale,
                               uint16_t* __restrict__ F, uint16_t* __restrict__ FT) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (i < n_valid) {
      put_features(i, n, X4[4 * (size_t)i], X4[4 * (size_t)i + 1], X4[4 * (size_t)i + 2], X4[4 * (size_t)i + 3],
                   r_scale, F, FT);
    } else {  // padding rows: a valid geometry, never read back
      put_features(i, n, 1.0, 0.0, 0.0, 0.0, r_scale, F, FT);
    }
  }
}

// Sample i = (radial point i / n_grid, grid point i % n_grid) of a PES grid:
// coords column-major (n_coord x n_grid), angles in rows ix, iy, iphi.
__global__ void k_mlp_grid_features(int n, int n_valid, int n_grid, const double* __restrict__ R, int n_coord,
                                    const double* __restrict__ coords, int ix, int iy, int iphi, double r_scale,
                                    uint16_t* __restrict__ F) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (i < n_valid) {
      const int g = i % n_grid;
      const double* c = coords + (size_t)g * n_coord;
      put_features(i, n, R[i / n_grid], c[ix], c[iy], c[iphi], r_scale, F, nullptr);
    } else {
      put_features(i, n, 1.0, 0.0, 0.0, 0.0, r_scale, F, nullptr);
    }
  }
}

// Mean-square loss of output column 0 against the targets over the first n_valid
// rows: loss += sum d^2 / n_valid; dZ[i][0] = 2 d / n_valid (BF16, row-major and
// transposed; the other columns stay zero); dbias[0] += sum 2 d / n_valid.
__global__ __launch_bounds__(256) void k_mlp_loss(int n, int n_valid, const float* __restrict__ Y,
                                                  const float* __restrict__ T, uint16_t* __restrict__ dZ,
                                                  uint16_t* __restrict__ dZT, float* __restrict__ dbias,
                                                  float* __restrict__ loss) {
  __shared__ float s_l[256], s_g[256];
  float l = 0.0f, gs = 0.0f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float g = 0.0f;
    if (i < n_valid) {
      const float d = Y[(size_t)i * 64] - T[i];
      l += d * d;
      g = 2.0f * d / (float)n_valid;
    }
    gs += g;
    const uint16_t h = to_bf16(g);
    dZ[(size_t)i * 64] = h;
    dZT[i] = h;
  }
  s_l[threadIdx.x] = l;
  s_g[threadIdx.x] = gs;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s_l[threadIdx.x] += s_l[threadIdx.x + s];
      s_g[threadIdx.x] += s_g[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(loss, s_l[0] / (float)n_valid);
    atomicAdd(dbias, s_g[0]);
  }
}

// out[c] += sum over rows of X[r][c] (X: M x N, N a multiple of 64); each
// workgroup sums a 64-column strip over its share of the rows.
__global__ __launch_bounds__(256) void k_mlp_colsum(int M, int N, const float* __restrict__ X, float* __restrict__ out) {
  __shared__ float s[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int g = threadIdx.x >> 6;
  float acc = 0.0f;
  for (int r = blockIdx.y * 4 + g; r < M; r += gridDim.y * 4) acc += X[(size_t)r * N + c];
  s[g][threadIdx.x & 63] = acc;
  __syncthreads();
  if (g == 0) atomicAdd(&out[c], s[0][threadIdx.x] + s[1][threadIdx.x] + s[2][threadIdx.x] + s[3][threadIdx.x]);
}

// Adam step on a rows x cols FP32 tensor; writes its BF16 copy and, when bfT is
// given, the BF16 transpose (cols x rows).
__global__ void k_mlp_adam(int rows, int cols, float* __restrict__ theta, const float* __restrict__ grad,
                           float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2, float eps,
                           float c1, float c2, uint16_t* __restrict__ bf, uint16_t* __restrict__ bfT) {
  const int n = rows * cols;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float t = theta[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    theta[i] = t;
    const uint16_t h = to_bf16(t);
    if (bf) bf[i] = h;
    if (bfT) bfT[(size_t)(i % cols) * rows + i / cols] = h;
  }
}

// BF16 copies of an FP32 tensor (rounding only), row-major and transposed.
__global__ void k_mlp_to_bf16(int rows, int cols, const float* __restrict__ x, uint16_t* __restrict__ bf,
                              uint16_t* __restrict__ bfT) {
  const int n = rows * cols;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint16_t h = to_bf16(x[i]);
    if (bf) bf[i] = h;
    if (bfT) bfT[(size_t)(i % cols) * rows + i / cols] = h;
  }
}

// Radial features of the angular-basis network, x = r_scale / R_i: columns 0-2
// hold x, x^2, x^3 and columns 3-5 the parts BF16 rounding drops from them,
// 256 (x^k - bf16(x^k)), so the network sees R to about 16 bits instead of the 8 of
// one BF16 input (zero beyond column 5; padding rows use x = 1/2).
__device__ __forceinline__ float bf16_value(float v) { return __uint_as_float((uint32_t)to_bf16(v) << 16); }

// lo[i] = bf16(x[i] - bf16(x[i])): the low BF16 part of an FP32 tensor.
__global__ void k_mlp_split_lo(int n, const float* __restrict__ x, uint16_t* __restrict__ lo) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    lo[i] = to_bf16(x[i] - bf16_value(x[i]));
}

__global__ void k_mlp_radial_features(int n, int n_valid, const double* __restrict__ R, double r_scale,
                                      uint16_t* __restrict__ F, uint16_t* __restrict__ FT) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double x = i < n_valid ? r_scale / R[i] : 0.5;
    const double pw[3] = {x, x * x, x * x * x};
    for (int c = 0; c < 64; ++c) {
      float v = 0.0f;
      if (c < 3) {
        v = (float)pw[c];
      } else if (c < 6) {
        const float p = (float)pw[c - 3];
        v = 256.0f * (p - bf16_value(p));
      }
      const uint16_t h = to_bf16(v);
      F[(size_t)i * 64 + c] = h;
      if (FT) FT[(size_t)c * n + i] = h;
    }
  }
}

// Loss of the angular-basis head.  Row i's prediction is the dot product of the
// network output U[i][:] (ld columns, FP32) with the angular functions of its
// sample, G[gidx[i]][:] (FP32, ld columns, zero past the basis): d = pred - T[i],
// loss += sum d^2 / n_valid, and the output gradient dZ[i][t] = 2 d G[gidx[i]][t] /
// n_valid -> dZf (FP32, for the bias gradient), dZ and dZT (BF16).  Rows past
// n_valid get zero gradients.  One wave per row.
__global__ __launch_bounds__(256) void k_mlp_basis_loss(int n, int n_valid, int ld, const float* __restrict__ U,
                                                        const int* __restrict__ gidx, const float* __restrict__ G,
                                                        const float* __restrict__ T, float* __restrict__ dZf,
                                                        uint16_t* __restrict__ dZ, uint16_t* __restrict__ dZT,
                                                        float* __restrict__ loss) {
  const int lane = threadIdx.x & 63;
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= n) return;
  const bool live = row < n_valid;
  const float* g = G + (size_t)(live ? gidx[row] : 0) * ld;
  float part = 0.0f;
  for (int t = lane; t < ld; t += 64) part += U[(size_t)row * ld + t] * g[t];
  for (int off = 32; off >= 1; off >>= 1) part += __shfl_xor(part, off, 64);
  const float d = live ? part - T[row] : 0.0f;
  const float c = 2.0f * d / (float)n_valid;
  for (int t = lane; t < ld; t += 64) {
    const float v = c * g[t];
    dZf[(size_t)row * ld + t] = v;
    const uint16_t h = to_bf16(v);
    dZ[(size_t)row * ld + t] = h;
    dZT[(size_t)t * n + row] = h;
  }
  if (lane == 0 && live) atomicAdd(loss, d * d / (float)n_valid);
}

// V[r][g] = scale[r] * sum_t U[r][t] GT[t][g] for the first nt of ld network
// outputs (U FP32, GT = the angular functions transposed, FP64): the network's
// PES on a grid of n_grid angular points at nR radial points.
__global__ __launch_bounds__(256) void k_mlp_basis_eval(int nR, int n_grid, int nt, int ld, const float* __restrict__ U,
                                                        const double* __restrict__ GT, const double* __restrict__ scale,
                                                        double* __restrict__ V) {
  __shared__ double u[512];
  const int r = blockIdx.y;
  for (int t = threadIdx.x; t < nt; t += blockDim.x) u[t] = (double)U[(size_t)r * ld + t];
  __syncthreads();
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_grid) return;
  double acc = 0.0;
  for (int t = 0; t < nt; ++t) acc = fma(u[t], GT[(size_t)t * n_grid + g], acc);
  V[(size_t)r * n_grid + g] = scale[r] * acc;
}

int blocks_for(long long n) {
  long long b = (n + 255) / 256;
  return (int)(b > 4096 ? 4096 : (b < 1 ? 1 : b));
}

}  // namespace

int mlp_gemm(int mode, int M, int N, int K, const void* A, const void* B, const float* bias, const float* Z, float* Cf,
             void* Cb, void* CbT, hipStream_t stream) {
  if (M <= 0 || N <= 0 || K <= 0 || M % 64 || N % 64 || K % 16)
    return fail(TICC_ERR_VALUE, "mlp_gemm: M, N must be positive multiples of 64 and K of 16 (got %d, %d, %d)", M, N, K);
  const auto* a = static_cast<const uint16_t*>(A);
  const auto* b = static_cast<const uint16_t*>(B);
  auto* cb = static_cast<uint16_t*>(Cb);
  auto* cbt = static_cast<uint16_t*>(CbT);
  const dim3 grid((M / 64) * (N / 64));
  switch (mode) {
    case EPI_STORE:
      hipLaunchKernelGGL(k_mlp_gemm<EPI_STORE>, grid, dim3(256), 0, stream, M, N, K, a, b, bias, Z, Cf, cb, cbt);
      break;
    case EPI_HIDDEN:
      hipLaunchKernelGGL(k_mlp_gemm<EPI_HIDDEN>, grid, dim3(256), 0, stream, M, N, K, a, b, bias, Z, Cf, cb, cbt);
      break;
    case EPI_OUTPUT:
      hipLaunchKernelGGL(k_mlp_gemm<EPI_OUTPUT>, grid, dim3(256), 0, stream, M, N, K, a, b, bias, Z, Cf, cb, cbt);
      break;
    case EPI_BACK:
      hipLaunchKernelGGL(k_mlp_gemm<EPI_BACK>, grid, dim3(256), 0, stream, M, N, K, a, b, bias, Z, Cf, cb, cbt);
      break;
    default:
      return fail(TICC_ERR_VALUE, "mlp_gemm: unknown epilogue %d", mode);
  }
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int mlp_gemm_split(int mode, int M, int N, int K, const void* A, const void* A_lo, const void* B, const void* B_lo,
                   const float* bias, float* Cf, void* Cb, void* Cb_lo, void* CbT, hipStream_t stream) {
  if (M <= 0 || N <= 0 || K <= 0 || M % 64 || N % 64 || K % 16)
    return fail(TICC_ERR_VALUE, "mlp_gemm_split: M, N must be positive multiples of 64 and K of 16 (got %d, %d, %d)", M,
                N, K);
  const auto* a = static_cast<const uint16_t*>(A);
  const auto* al = static_cast<const uint16_t*>(A_lo);
  const auto* b = static_cast<const uint16_t*>(B);
  const auto* bl = static_cast<const uint16_t*>(B_lo);
  const dim3 grid((M / 64) * (N / 64));
  if (mode == EPI_HIDDEN)
    hipLaunchKernelGGL((k_mlp_gemm<EPI_HIDDEN, true>), grid, dim3(256), 0, stream, M, N, K, a, b, bias, nullptr, Cf,
                       static_cast<uint16_t*>(Cb), static_cast<uint16_t*>(CbT), al, bl, static_cast<uint16_t*>(Cb_lo));
  else if (mode == EPI_OUTPUT)
    hipLaunchKernelGGL((k_mlp_gemm<EPI_OUTPUT, true>), grid, dim3(256), 0, stream, M, N, K, a, b, bias, nullptr, Cf,
                       nullptr, nullptr, al, bl, nullptr);
  else
    return fail(TICC_ERR_VALUE, "mlp_gemm_split: mode must be 1 (hidden) or 2 (output), got %d", mode);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int mlp_split_lo(int n, const float* x, void* lo, hipStream_t stream) {
  hipLaunchKernelGGL(k_mlp_split_lo, dim3(blocks_for(n)), dim3(256), 0, stream, n, x, static_cast<uint16_t*>(lo));
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int mlp_features(int n, int n_valid, const double* X4, double r_scale, void* F, void* FT, hipStream_t stream) {
  hipLaunchKernelGGL(k_mlp_features, dim3(blocks_for(n)), dim3(256), 0, stream, n, n_valid, X4, r_scale,
                     static_cast<uint16_t*>(F), static_cast<uint16_t*>(FT));
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int mlp_grid_features(int n, int n_valid, int n_grid, const double* R, int n_coord, const double* coords, int ix, int iy,
                      int iphi, double r_scale, void* F, hipStream_t stream) {
  hipLaunchKernelGGL(k_mlp_grid_features, dim3(blocks_for(n)), dim3(256), 0, stream, n, n_valid, n_grid, R, n_coord,
                     coords, ix, iy, iphi, r_scale, static_cast<uint16_t*>(F));
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int mlp_loss(int n, int n_valid, const float* Y, const float* T, void* dZ, void* dZT, float* dbias, float* loss,
             hipStream_t stream) {
  hipLaunchKernelGGL(k_mlp_loss, dim3(blocks_for(n) < 1024 ? blocks_for(n) : 1024), dim3(256), 0, stream, n, n_valid,
                     Y, T, static_cast<uint16_t*>(dZ), static_cast<uint16_t*>(dZT), dbias, loss);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int mlp_radial_features(int n, int n_valid, const double* R, double r_scale, void* F, void* FT, hipStream_t stream) {
  hipLaunchKernelGGL(k_mlp_radial_features, dim3(blocks_for(n)), dim3(256), 0, stream, n, n_valid, R, r_scale,
                     static_cast<uint16_t*>(F), static_cast<uint16_t*>(FT));
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int mlp_basis_loss(int n, int n_valid, int ld, const float* U, const int* gidx, const float* G, const float* T,
                   float* dZf, void* dZ, void* dZT, float* loss, hipStream_t stream) {
  hipLaunchKernelGGL(k_mlp_basis_loss, dim3((n + 3) / 4), dim3(256), 0, stream, n, n_valid, ld, U, gidx, G, T, dZf,
                     static_cast<uint16_t*>(dZ), static_cast<uint16_t*>(dZT), loss);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int mlp_basis_eval(int nR, int n_grid, int nt, int ld, const float* U, const double* GT, const double* scale, double* V,
                   hipStream_t stream) {
  hipLaunchKernelGGL(k_mlp_basis_eval, dim3((n_grid + 255) / 256, nR), dim3(256), 0, stream, nR, n_grid, nt, ld, U, GT,
                     scale, V);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int mlp_colsum(int M, int N, const float* X, float* out, hipStream_t stream) {
  if (N % 64) return fail(TICC_ERR_VALUE, "mlp_colsum: N must be a multiple of 64");
  int gy = (M + 4 * 64 - 1) / (4 * 64);
  if (gy > 256) gy = 256;
  hipLaunchKernelGGL(k_mlp_colsum, dim3(N / 64, gy), dim3(256), 0, stream, M, N, X, out);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int mlp_adam(int rows, int cols, float* theta, const float* grad, float* m, float* v, float lr, float b1, float b2,
             float eps, int step, void* bf, void* bfT, hipStream_t stream) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for((long long)rows * cols)), dim3(256), 0, stream, rows, cols, theta, grad,
                     m, v, lr, b1, b2, eps, c1, c2, static_cast<uint16_t*>(bf), static_cast<uint16_t*>(bfT));
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int mlp_to_bf16(int rows, int cols, const float* x, void* bf, void* bfT, hipStream_t stream) {
  hipLaunchKernelGGL(k_mlp_to_bf16, dim3(blocks_for((long long)rows * cols)), dim3(256), 0, stream, rows, cols, x,
                     static_cast<uint16_t*>(bf), static_cast<uint16_t*>(bfT));
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

}  // namespace ticc
