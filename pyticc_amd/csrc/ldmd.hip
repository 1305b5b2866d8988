// Manolopoulos LDMD sweep (kernel B) for an energy batch.
//
// Per sector s with half step h (all matrices FP64, padded to Np):
//   ref    = diag(W_mid)                          (energy independent)
//   Q_s    = h (W_start - diag ref) / 3,  Q_e = h (W_end - diag ref) / 3
//   Q_m    = (4/h) [ (I - h^2 (W_mid - diag ref)/6)^-1 - I ]   (batched over sectors)
//   per energy e:  p^2 = ref - 2 mu E_e -> y1, y2 (coth/cot, csch/csc or series)
//     A1 = Y + Q_s + diag y1                 -> X1 = A1^-1   (blocked GJ)
//     Ym = Q_m + diag y1 - y2 X1 y2
//     A2 = Ym + Q_m + diag y1                -> X2 = A2^-1
//     Ye = Q_e + diag y1 - y2 X2 y2,   Y <- (Ye + Ye^T) / 2
// Q_s and Q_e are never stored: the form kernels read W_start / W_end directly.
// The elementwise passes fuse the column unpermute of the GJ result.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "ticc_common.h"

namespace ticc {
namespace {

constexpr int QCHUNK = 32;     // sectors whose Q_mid are inverted in one batch
constexpr int MAX_GROUPS = 4;  // energy groups swept on separate streams

__device__ __forceinline__ void reference_values(double h, double p2, double* y1, double* y2) {
  const double mag = sqrt(fabs(p2));
  const double arg = mag * h;
  if (fabs(arg) < 1.0e-5) {
    *y1 = 1.0 / h + p2 * h / 3.0 - p2 * p2 * h * h * h / 45.0;
    *y2 = 1.0 / h - p2 * h / 6.0 + 7.0 * (p2 * p2) * (h * h * h) / 360.0;
  } else if (p2 >= 0.0) {
    *y1 = mag / tanh(arg);
    *y2 = mag / sinh(arg);
  } else {
    *y1 = mag / tan(arg);
    *y2 = mag / sin(arg);
  }
}

// Blocks: batch item item0 + e belongs to block (item0 + e) / per, whose W_base
// lies gstride elements after the previous block's (gstride = 0: one block).
__device__ __forceinline__ size_t block_off(int item0, int e, int per, long long gstride) {
  return (size_t)((item0 + e) / per) * (size_t)gstride;
}

__global__ void k_refvals(double h, const double* __restrict__ Wm, int N, int Np, const double* __restrict__ E,
                          double two_mu, int nE, double* __restrict__ y1, double* __restrict__ y2, int item0, int per,
                          long long gstride) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nE * Np) return;
  const int e = idx / Np;
  const int i = idx % Np;
  double a = 0.0, b = 0.0;
  if (i < N) {
    const double p2 = Wm[block_off(item0, e, per, gstride) + (size_t)i * N + i] - two_mu * E[e];
    reference_values(h, p2, &a, &b);
  }
  y1[idx] = a;
  y2[idx] = b;
}

// A1 = (Y + Q_s) + diag y1, identity padding.
__global__ void k_form1(const double* __restrict__ Y, const double* __restrict__ Ws0, const double* __restrict__ Wm0,
                        double h, const double* __restrict__ y1, int N, int Np, double* __restrict__ A, int item0, int per,
                        long long gstride) {
  const int e = blockIdx.y;
  const size_t n2 = (size_t)Np * Np;
  const double* Ws = Ws0 + block_off(item0, e, per, gstride);
  const double* Wm = Wm0 + block_off(item0, e, per, gstride);
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n2; idx += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / Np);
    const int j = (int)(idx % Np);
    double v;
    if (i < N && j < N) {
      double t = Ws[(size_t)i * N + j];
      if (i == j) t = t - Wm[(size_t)i * N + i];
      const double qs = h * t / 3.0;
      v = Y[e * n2 + idx] + qs;
      if (i == j) v += y1[(size_t)e * Np + i];
    } else {
      v = (i == j) ? 1.0 : 0.0;
    }
    A[e * n2 + idx] = v;
  }
}

// A2 = (Ym + Q_m) + diag y1,  Ym = (Q_m + diag y1) - y2 (X1 y2).
// One wave per row (4 rows per 256-thread workgroup), lanes along the row.
__global__ __launch_bounds__(256) void k_form2(const double* __restrict__ X1, const int* __restrict__ pinv,
                                               const double* __restrict__ Qm0, const double* __restrict__ y1,
                                               const double* __restrict__ y2, int N, int Np, double* __restrict__ A,
                                               int item0, int per, long long qstride) {
  const int e = blockIdx.y;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= Np) return;
  const int lane = threadIdx.x & 63;
  const size_t n2 = (size_t)Np * Np;
  const double* Qr = Qm0 + block_off(item0, e, per, qstride) + (size_t)i * Np;
  const double* Xr = X1 + e * n2 + (size_t)i * Np;
  const int* pe = pinv + (size_t)e * Np;
  const double* y2e = y2 + (size_t)e * Np;
  double* Ar = A + e * n2 + (size_t)i * Np;
  const double y1i = (i < N) ? y1[(size_t)e * Np + i] : 0.0;
  const double y2i = (i < N) ? y2e[i] : 0.0;
  for (int j = lane; j < Np; j += 64) {
    double v;
    if (i < N && j < N) {
      const double x = Xr[pe[j]];
      const double q = Qr[j];
      double ym = q;
      if (i == j) ym += y1i;
      ym = ym - y2i * (x * y2e[j]);
      v = ym + q;
      if (i == j) v += y1i;
    } else {
      v = (i == j) ? 1.0 : 0.0;
    }
    Ar[j] = v;
  }
}

// Y = (Ye + Ye^T)/2,  Ye = (Q_e + diag y1) - y2 (X2 y2).  32 x 32 tiles through LDS.
// With half > 0 the matrices are the real form [[Re Z, -Im Z], [Im Z, Re Z]]
// of complex symmetric N x N matrices (N = half): Z -> (Z + Z^T)/2 is then
// (Ye + P Ye^T P)/2 with P = diag(I, -I), i.e. the transposed element enters
// with a minus sign when exactly one of i, j lies in the second half.
// (we = W_end[i][j], wm = W_mid[i][i], read by the caller)
__device__ __forceinline__ double yend_elem(const double* __restrict__ X2e, const int* __restrict__ pinve, double we,
                                            double wm, double h, const double* __restrict__ y1e,
                                            const double* __restrict__ y2e, int Np, int i, int j) {
  double t = we;
  if (i == j) t = t - wm;
  double ye = h * t / 3.0;
  if (i == j) ye += y1e[i];
  const double x = X2e[(size_t)i * Np + pinve[j]];
  return ye - y2e[i] * (x * y2e[j]);
}

// The next sector's A1 = (Y + Q_s) + diag y1 from element (i, j) of Y: the
// arithmetic of k_form1, with the addition rounded on its own as when Y is
// stored and read back.
struct NextSector {
  const double* Ws;  // W_start and W_mid of the next sector (block 0)
  const double* Wm;
  double h;
  const double* y1;  // its y1 (batch x Np)
  double* A;         // nullptr: write Y instead
};

// (ws = the next sector's W_start[i][j], wm = its W_mid[i][i], read by the caller)
__device__ __forceinline__ double next_a1(const NextSector& n, double ws, double wm, const double* __restrict__ y1e,
                                          int N, int i, int j, double v) {
#pragma clang fp contract(off)
  if (i >= N || j >= N) return (i == j) ? 1.0 : 0.0;
  double t = ws;
  if (i == j) t = t - wm;
  const double qs = n.h * t / 3.0;
  double a = v + qs;
  if (i == j) a += y1e[i];
  return a;
}

// One workgroup per pair of mirror 32 x 32 tiles (ti <= tj) and energy: both tiles
// of Ye are formed once into LDS and both tiles of Y written, so X2 is read once.
// With nx.A set, the next sector's A1 is written instead of Y (k_form1 fused).
// (Energy-major grids and groups of energies per workgroup sharing register-held
// W tiles were measured and were not faster: 280 and 500 us against 269 us at
// N = 800 x 64.)
__global__ __launch_bounds__(256) void k_form3(const double* __restrict__ X2, const int* __restrict__ pinv,
                                               const double* __restrict__ We0, const double* __restrict__ Wm0,
                                               double h, const double* __restrict__ y1, const double* __restrict__ y2,
                                               int N, int Np, int half, double* __restrict__ Y, int item0, int per,
                                               long long gstride, NextSector nx) {
  __shared__ double E1[32][33];
  __shared__ double E2[32][33];
  const int e = blockIdx.y;
  const size_t boff = block_off(item0, e, per, gstride);
  const double* We = We0 + boff;
  const double* Wm = Wm0 + boff;
  const double* nWs = nx.Ws + (nx.A ? boff : 0);
  const double* nWm = nx.Wm + (nx.A ? boff : 0);
  const double* ny1 = nx.y1 + (size_t)e * Np;
  double* Out = nx.A ? nx.A : Y;
  // pair index -> (ti, tj), 0 <= ti <= tj
  const int pidx = blockIdx.x;
  int tj = (int)((sqrt(8.0 * pidx + 1.0) - 1.0) * 0.5);
  while (tj * (tj + 1) / 2 > pidx) --tj;
  while ((tj + 1) * (tj + 2) / 2 <= pidx) ++tj;
  const int ti = pidx - tj * (tj + 1) / 2;
  const bool mirror = ti != tj;
  const size_t n2 = (size_t)Np * Np;
  const double* X2e = X2 + e * n2;
  const int* pinve = pinv + (size_t)e * Np;
  const double* y1e = y1 + (size_t)e * Np;
  const double* y2e = y2 + (size_t)e * Np;
  const int tx = threadIdx.x & 31;
  const int ty = threadIdx.x >> 5;  // 0..7
  for (int r = ty; r < 32; r += 8) {
    const int i = ti * 32 + r, j = tj * 32 + tx;
    E1[r][tx] = (i < N && j < N)
                    ? yend_elem(X2e, pinve, We[(size_t)i * N + j], (i == j) ? Wm[(size_t)i * N + i] : 0.0, h, y1e,
                                y2e, Np, i, j)
                    : 0.0;
    if (mirror) {
      const int i2 = tj * 32 + r, j2 = ti * 32 + tx;
      E2[r][tx] = (i2 < N && j2 < N)
                      ? yend_elem(X2e, pinve, We[(size_t)i2 * N + j2], (i2 == j2) ? Wm[(size_t)i2 * N + i2] : 0.0, h,
                                  y1e, y2e, Np, i2, j2)
                      : 0.0;
    }
  }
  __syncthreads();
  double (*Et)[33] = mirror ? E2 : E1;  // the mirror tile (tj, ti)
  for (int r = ty; r < 32; r += 8) {
    const int i = ti * 32 + r, j = tj * 32 + tx;
    double v = 0.0;
    if (i < N && j < N) {
      const double a = E1[r][tx], t = Et[tx][r];
      const bool flip = half > 0 && ((i < half) != (j < half));
      v = 0.5 * (flip ? a - t : a + t);
    }
    Out[e * n2 + (size_t)i * Np + j] =
        nx.A ? next_a1(nx, (i < N && j < N) ? nWs[(size_t)i * N + j] : 0.0,
                       (i < N && i == j) ? nWm[(size_t)i * N + i] : 0.0, ny1, N, i, j, v)
             : v;
    if (mirror) {
      const int i2 = tj * 32 + r, j2 = ti * 32 + tx;
      double v2 = 0.0;
      if (i2 < N && j2 < N) {
        const double a = E2[r][tx], t = E1[tx][r];
        const bool flip = half > 0 && ((i2 < half) != (j2 < half));
        v2 = 0.5 * (flip ? a - t : a + t);
      }
      Out[e * n2 + (size_t)i2 * Np + j2] =
          nx.A ? next_a1(nx, (i2 < N && j2 < N) ? nWs[(size_t)i2 * N + j2] : 0.0,
                         (i2 < N && i2 == j2) ? nWm[(size_t)i2 * N + i2] : 0.0, ny1, N, i2, j2, v2)
               : v2;
    }
  }
}

// M_s = I - (h_s^2 T_mid)/6 for a chunk of sectors.
// (grid.y = n_block * nc: item blockIdx.y = block * nc + sector)
__global__ void k_qm_form(const double* __restrict__ Wm0, long long stride, const double* __restrict__ hs, int N,
                          int Np, double* __restrict__ M, int nc, long long gstride) {
  const int s = blockIdx.y % nc;
  const int g = blockIdx.y / nc;
  const double h = hs[s];
  const double* Wm = Wm0 + (size_t)g * gstride + s * stride;
  const size_t n2 = (size_t)Np * Np;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n2; idx += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / Np);
    const int j = (int)(idx % Np);
    double v;
    if (i < N && j < N) {
      double t = Wm[(size_t)i * N + j];
      if (i == j) t = t - Wm[(size_t)i * N + i];
      v = ((i == j) ? 1.0 : 0.0) - (h * h) * t / 6.0;
    } else {
      v = (i == j) ? 1.0 : 0.0;
    }
    M[blockIdx.y * n2 + idx] = v;
  }
}

// Q_m = (4/h) (M^-1 - I).
// (grid.y = n_block * nc as k_qm_form; block g's sector s goes to Qm[g * qcap + s])
__global__ void k_qm_finish(const double* __restrict__ Xi, const int* __restrict__ pinv, const double* __restrict__ hs,
                            int N, int Np, double* __restrict__ Qm, int nc, int qcap) {
  const int b = blockIdx.y;
  const int s = b % nc;
  const int g = b / nc;
  const double c = 4.0 / hs[s];
  const size_t n2 = (size_t)Np * Np;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n2; idx += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / Np);
    const int j = (int)(idx % Np);
    double v = 0.0;
    if (i < N && j < N) v = c * (Xi[b * n2 + (size_t)i * Np + pinv[(size_t)b * Np + j]] - ((i == j) ? 1.0 : 0.0));
    Qm[((size_t)g * qcap + s) * n2 + idx] = v;
  }
}

__global__ void k_init_inelastic(const double* __restrict__ W0, const double* __restrict__ E, double two_mu, int N,
                                 int nE, double* __restrict__ Y) {
  const size_t n2 = (size_t)N * N;
  const size_t total = n2 * nE;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const int e = (int)(idx / n2);
    const size_t rem = idx % n2;
    const int i = (int)(rem / N);
    const int j = (int)(rem % N);
    double v = 0.0;
    if (i == j) v = sqrt(fabs(W0[(size_t)i * N + i] - two_mu * E[e]));
    Y[idx] = v;
  }
}

// Capture boundary condition in real form: Z_e = diag(z), z_i = sqrt(d) for
// d = W0_ii - 2 mu E_e >= 0 and -i sqrt(-d) otherwise; Y_e = [[Re Z, -Im Z],
// [Im Z, Re Z]] of size 2N.
__global__ void k_init_capture(const double* __restrict__ W0, const double* __restrict__ E, double two_mu, int N,
                               int nE, double* __restrict__ Y) {
  const int N2 = 2 * N;
  const size_t n2 = (size_t)N2 * N2;
  const size_t total = n2 * nE;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const int e = (int)(idx / n2);
    const size_t rem = idx % n2;
    const int i = (int)(rem / N2);
    const int j = (int)(rem % N2);
    double v = 0.0;
    if (i % N == j % N) {
      const int c = i % N;
      const double d = W0[(size_t)c * N + c] - two_mu * E[e];
      const double re = d >= 0.0 ? sqrt(d) : 0.0;
      const double im = d < 0.0 ? -sqrt(-d) : 0.0;
      const bool top = i < N, left = j < N;
      v = (top == left) ? re : (top ? -im : im);
    }
    Y[idx] = v;
  }
}

// FP32 copy of an FP64 batch (the mixed sweep's input to the FP32 inversion).
__global__ void k_to_f32(const double* __restrict__ in, float* __restrict__ out, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}

__global__ void k_fill_u32(unsigned* __restrict__ p, int n, unsigned v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// mask[b] = 1 where the FP32 inverse's residual max|I - A X| exceeds tau (its
// Newton steps would not reach FP64 accuracy): that matrix is inverted in FP64.
__global__ void k_flag(const unsigned* __restrict__ rho, int n, unsigned tau_bits, int* __restrict__ mask,
                       unsigned long long* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int f = rho[i] > tau_bits ? 1 : 0;
  mask[i] = f;
  if (f) atomicAdd(count, 1ull);
}

// The identity column map (a plain inverse needs no unpermute).
__global__ void k_iota(int* __restrict__ p, int Np, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = (int)(i % Np);
}

// Counters of the calling thread's mixed-precision sweeps (see mixed_stats).
thread_local long long t_mixed_inversions = 0;
thread_local long long t_mixed_fallbacks = 0;

int grid1(long long n) {
  long long b = (n + 255) / 256;
  if (b > 2048) b = 2048;
  return (int)(b < 1 ? 1 : b);
}

}  // namespace

int mixed_stats(long long* inversions, long long* fallbacks, int reset) {
  *inversions = t_mixed_inversions;
  *fallbacks = t_mixed_fallbacks;
  if (reset) t_mixed_inversions = t_mixed_fallbacks = 0;
  return TICC_OK;
}

int to_f32(const double* in, float* out, long long n, hipStream_t stream) {
  hipLaunchKernelGGL(k_to_f32, dim3(grid1(n)), dim3(256), 0, stream, in, out, n);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int ldmd_init(int nE, int N, const double* W0, const double* E_host, double mu, double* Y, hipStream_t stream) {
  void* ws = nullptr;
  TICC_RETURN_IF(workspace(WS_MATCH, sizeof(double) * nE, &ws));
  double* dE = static_cast<double*>(ws);
  TICC_CHECK_HIP(hipMemcpyAsync(dE, E_host, sizeof(double) * nE, hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_init_inelastic, dim3(grid1((long long)nE * N * N)), dim3(256), 0, stream, W0, dE, 2.0 * mu, N,
                     nE, Y);
  TICC_CHECK_LAUNCH();
  TICC_CHECK_HIP(hipStreamSynchronize(stream));
  return TICC_OK;
}

int ldmd_init_capture(int nE, int N, const double* W0, const double* E_host, double mu, double* Y, hipStream_t stream) {
  void* ws = nullptr;
  TICC_RETURN_IF(workspace(WS_MATCH, sizeof(double) * nE, &ws));
  double* dE = static_cast<double*>(ws);
  TICC_CHECK_HIP(hipMemcpyAsync(dE, E_host, sizeof(double) * nE, hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_init_capture, dim3(grid1((long long)nE * 4 * N * N)), dim3(256), 0, stream, W0, dE, 2.0 * mu,
                     N, nE, Y);
  TICC_CHECK_LAUNCH();
  TICC_CHECK_HIP(hipStreamSynchronize(stream));
  return TICC_OK;
}

int ldmd_propagate(int nE, int N, const double* E_host, double mu, int nS, const double* hs_host, const double* Ws0,
                   const double* Wm0, const double* We0, long long wstride, double* Y, hipStream_t stream, int half,
                   int n_block, long long gstride, int refine, SweepAsync* as) {
  const int Np = round_up(N, GJ_NB);
  if (Np > MAX_NP) return fail(TICC_ERR_UNSUPPORTED, "n_channel=%d exceeds the supported maximum %d", N, MAX_NP);
  if (n_block < 1 || nE % n_block) return fail(TICC_ERR_VALUE, "the batch must hold the same energies for every block");
  // batch item b belongs to block b / per; each block's Q_mid chunk holds qc sectors
  const int per_block = nE / n_block;
  const int qc = n_block == 1 ? QCHUNK : (QCHUNK / n_block > 0 ? QCHUNK / n_block : 1);
  // Energy groups run on their own streams: with two, one group's serial pieces
  // (its first panel chain, look-ahead updates, form kernels) run beside the
  // other group's fused launches, 3-5% faster at N = 800 and 400 with 64
  // energies (4.5% for the mixed-precision sweep), results bit for bit those
  // of one group (three groups are 35% slower at N = 800).  Two groups are the
  // default of the eager inelastic sweep of one block with at least 16
  // energies and Np >= 256; TICC_SWEEP_GROUPS overrides it.
  const bool mixed = refine >= 0;
  int G = (!as && half == 0 && n_block == 1 && nE >= 16 && Np >= 256) ? 2 : 1;
  if (const char* env = getenv("TICC_SWEEP_GROUPS")) {  // tuning knob: 1..MAX_GROUPS
    const int v = atoi(env);
    if (v >= 1 && v <= MAX_GROUPS) G = v < nE ? v : nE;
  }
  const int per = (nE + G - 1) / G;
  const size_t n2 = (size_t)Np * Np;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t oY = take(sizeof(double) * nE * n2);
  const size_t oQ = take(sizeof(double) * 2 * n_block * qc * n2);
  const size_t oM = take(sizeof(double) * 2 * n_block * qc * n2);
  const size_t oE = take(sizeof(double) * nE);
  const size_t oH = take(sizeof(double) * nS);
  const size_t oFlag = take(sizeof(int));
  const size_t oCount = take(sizeof(unsigned long long));  // mixed sweep: matrices inverted in FP64
  // refine >= 0: the mixed-precision sweep -- each sector matrix is inverted
  // in FP32 and refined by `refine` Newton steps X <- X + X (I - A X), the
  // residual I - A X on the FP64 MFMA from the FP64 A, the correction X (I - A X)
  // on the FP32 MFMA; everything else (forms, reference values, Q_mid, Y) stays
  // FP64.  refine < 0: the FP64 sweep.
  struct Scratch {
    size_t A0, A1, P, src, piv, perm, pinv, y1, y2;
    size_t F0, F1, X, T32, X32[2], ipinv, XB, rho, mask;  // mixed sweep only
  };
  auto scratch = [&](int batch, bool yv) {
    Scratch sc{};
    if (yv && mixed) {
      sc.F0 = take(sizeof(float) * batch * n2);
      sc.F1 = take(sizeof(float) * batch * n2);
      sc.X = take(sizeof(double) * batch * n2);
      sc.T32 = take(sizeof(float) * batch * n2);
      sc.X32[0] = take(sizeof(float) * batch * n2);
      sc.X32[1] = take(sizeof(float) * batch * n2);
      sc.ipinv = take(sizeof(int) * batch * Np);
      sc.XB = take(sizeof(double) * batch * n2);
      sc.rho = take(sizeof(unsigned) * batch);
      sc.mask = take(sizeof(int) * batch);
    }
    sc.A0 = yv ? take(sizeof(double) * batch * n2) : 0;
    sc.A1 = yv ? take(sizeof(double) * batch * n2) : 0;
    sc.P = take(sizeof(double) * GJ_POUT_PANELS * batch * Np * GJ_NB);
    sc.src = take(sizeof(int) * GJ_SRC_MAPS * batch * Np);
    sc.piv = take(sizeof(int) * GJ_PIV_LISTS * batch * GJ_NB);
    sc.perm = take(sizeof(int) * batch * Np);
    sc.pinv = take(sizeof(int) * batch * Np);
    // y1, y2 of two sectors: the next sector's are formed before this one's last pass
    sc.y1 = yv ? take(sizeof(double) * 2 * batch * Np) : 0;
    sc.y2 = yv ? take(sizeof(double) * 2 * batch * Np) : 0;
    return sc;
  };
  const Scratch sm = scratch(n_block * qc, false);
  Scratch sg[MAX_GROUPS];
  for (int g = 0; g < G; ++g) sg[g] = scratch(per, true);
  void* base = nullptr;
  if (as) {  // enqueue-only form: caller-owned scratch, device-resident inputs
    if (refine >= 0 || half > 0) return fail(TICC_ERR_UNSUPPORTED, "the enqueue-only sweep is the FP64 inelastic sweep");
    if (!as->scratch) {
      as->scratch_bytes = off;
      return TICC_OK;
    }
    if (as->scratch_bytes < off)
      return fail(TICC_ERR_VALUE, "sweep scratch of %zu bytes is below the %zu bytes required", as->scratch_bytes, off);
    base = as->scratch;
  } else {
    TICC_RETURN_IF(workspace(WS_SWEEP, off, &base));
  }
  char* w = static_cast<char*>(base);
  auto D = [&](size_t o) { return reinterpret_cast<double*>(w + o); };
  auto I = [&](size_t o) { return reinterpret_cast<int*>(w + o); };
  auto F = [&](size_t o) { return reinterpret_cast<float*>(w + o); };
  double* Yp = D(oY);
  double* Qm = D(oQ);
  double* M0 = D(oM);
  double* M1 = M0 + (size_t)n_block * qc * n2;
  const double* dE = as ? as->E_dev : D(oE);
  const double* dH = as ? as->hs_dev : D(oH);
  int* flag = I(oFlag);

  hipStream_t gs[MAX_GROUPS];
  hipEvent_t ev[2 + 2 * MAX_GROUPS];
  TICC_RETURN_IF(aux_streams(G, gs));
  TICC_RETURN_IF(aux_events(2 + 2 * G, ev));
  hipEvent_t* ev_qm = ev;             // [2] Q_mid chunk ready (per buffer)
  hipEvent_t* ev_done = ev + 2;       // [G][2] group finished a chunk (per buffer)

  if (!as) {
    TICC_CHECK_HIP(hipMemcpyAsync(D(oE), E_host, sizeof(double) * nE, hipMemcpyHostToDevice, stream));
    TICC_CHECK_HIP(hipMemcpyAsync(D(oH), hs_host, sizeof(double) * nS, hipMemcpyHostToDevice, stream));
    TICC_CHECK_HIP(hipMemsetAsync(flag, 0, sizeof(int), stream));
  }
  unsigned long long* fb_count = reinterpret_cast<unsigned long long*>(w + oCount);
  if (mixed) TICC_CHECK_HIP(hipMemsetAsync(fb_count, 0, sizeof(unsigned long long), stream));
  float tau = 1.0e-3f;  // residual bound for keeping the refined FP32 inverse
  if (const char* env = getenv("TICC_MIXED_FALLBACK_TAU")) tau = (float)atof(env);  // tuning knob
  unsigned tau_bits = 0;
  memcpy(&tau_bits, &tau, sizeof(tau));
  TICC_RETURN_IF(pad_identity(as ? as->Y_in : Y, (long long)N * N, N, Yp, Np, nE, stream));
  if (mixed)
    for (int g = 0; g < G; ++g) {
      hipLaunchKernelGGL(k_iota, dim3(grid1((long long)per * Np)), dim3(256), 0, stream, I(sg[g].ipinv), Np,
                         (long long)per * Np);
      TICC_CHECK_LAUNCH();
    }

  const double two_mu = 2.0 * mu;
  const int gelem = grid1((long long)n2);
  // per group: the buffer holding the current sector's A1 and its GJ partner
  double* a_in[MAX_GROUPS];
  double* a_alt[MAX_GROUPS];
  const int tn = Np / 32;
  const int n_chunk = (nS + qc - 1) / qc;
  const long long qstride = (long long)qc * n2;  // one block's Q_mid chunk
  for (int ci = 0; ci < n_chunk; ++ci) {
    const int s0 = ci * qc;
    const int nc = (nS - s0) < qc ? (nS - s0) : qc;
    const int qb = ci & 1;
    double* Qbuf = Qm + (size_t)qb * n_block * qc * n2;
    if (ci >= 2)  // the groups are done with the chunk that last used this buffer
      for (int g = 0; g < G; ++g) TICC_CHECK_HIP(hipStreamWaitEvent(stream, ev_done[2 * g + qb], 0));
    hipLaunchKernelGGL(k_qm_form, dim3(gelem, n_block * nc), dim3(256), 0, stream, Wm0 + s0 * wstride, wstride, dH + s0,
                       N, Np, M0, nc, gstride);
    TICC_CHECK_LAUNCH();
    double* res = nullptr;
    TICC_RETURN_IF(
        gj_invert_raw(M0, M1, n_block * nc, Np, &res, I(sm.perm), I(sm.pinv), D(sm.P), I(sm.src), I(sm.piv), stream));
    hipLaunchKernelGGL(k_qm_finish, dim3(gelem, n_block * nc), dim3(256), 0, stream, res, I(sm.pinv), dH + s0, N, Np,
                       Qbuf, nc, qc);
    TICC_CHECK_LAUNCH();
    TICC_CHECK_HIP(hipEventRecord(ev_qm[qb], stream));

    for (int g = 0; g < G; ++g) {
      const int e0 = g * per;
      const int ne = (nE - e0) < per ? (nE - e0) : per;
      if (ne <= 0) continue;
      hipStream_t st = gs[g];
      const Scratch& sc = sg[g];
      double* A0 = D(sc.A0);
      double* A1 = D(sc.A1);
      double* Yg = Yp + (size_t)e0 * n2;
      auto ys = [&](size_t o, int s) { return D(o) + (size_t)(s & 1) * per * Np; };
      auto refvals = [&](int s) {
        hipLaunchKernelGGL(k_refvals, dim3((ne * Np + 255) / 256), dim3(256), 0, st, hs_host[s], Wm0 + s * wstride, N,
                           Np, dE + e0, two_mu, ne, ys(sc.y1, s), ys(sc.y2, s), e0, per_block, gstride);
        return hipGetLastError();
      };
      // Inverse of the sector matrices in Ain (Aalt: the FP64 inversion's ping-pong
      // partner): *X receives the buffer holding it and *pv its column map.
      auto invert = [&](double* Ain, double* Aalt, double** X, const int** pv) -> int {
        if (!mixed) {
          TICC_RETURN_IF(gj_invert_raw(Ain, Aalt, ne, Np, X, I(sc.perm), I(sc.pinv), D(sc.P), I(sc.src), I(sc.piv), st));
          *pv = I(sc.pinv);
          return TICC_OK;
        }
        const long long nn = (long long)ne * n2;
        hipLaunchKernelGGL(k_to_f32, dim3(grid1(nn)), dim3(256), 0, st, Ain, F(sc.F0), nn);
        TICC_CHECK_LAUNCH();
        float* r32 = nullptr;
        TICC_RETURN_IF(gj_invert_raw_f32(F(sc.F0), F(sc.F1), ne, Np, &r32, I(sc.perm), I(sc.pinv),
                                         reinterpret_cast<float*>(D(sc.P)), I(sc.src), I(sc.piv), st));
        TICC_RETURN_IF(gj_unpermute_f32(r32, D(sc.X), ne, Np, I(sc.pinv), st));
        if (refine > 0) TICC_RETURN_IF(gj_unpermute_f32f32(r32, F(sc.X32[0]), ne, Np, I(sc.pinv), st));
        unsigned* rho = reinterpret_cast<unsigned*>(w + sc.rho);
        int* mask = I(sc.mask);
        for (int it = 0; it < refine; ++it) {
          // T = I - A X (FP64 operands, FP32 result), X <- X + X T (FP32 operands);
          // the first residual's largest element decides which matrices fall back
          if (it == 0) {
            hipLaunchKernelGGL(k_fill_u32, dim3((ne + 255) / 256), dim3(256), 0, st, rho, ne, 0u);
            TICC_CHECK_LAUNCH();
          }
          TICC_RETURN_IF(refine_gemm_f64(Np, ne, Ain, D(sc.X), -1.0, nullptr, 0.0, 1.0, nullptr, F(sc.T32),
                                         it == 0 ? rho : nullptr, st));
          if (it == 0) {
            hipLaunchKernelGGL(k_flag, dim3((ne + 255) / 256), dim3(256), 0, st, rho, ne, tau_bits, mask, fb_count);
            TICC_CHECK_LAUNCH();
          }
          TICC_RETURN_IF(refine_gemm_f32(Np, ne, F(sc.X32[it & 1]), F(sc.T32), 1.0, D(sc.X), 1.0, 0.0, D(sc.X),
                                         it + 1 < refine ? F(sc.X32[(it + 1) & 1]) : nullptr, st));
        }
        if (refine > 0) {  // the flagged matrices: FP64 inversion of A in place (every kernel skips the others)
          double* rf = nullptr;
          TICC_RETURN_IF(gj_invert_raw_masked(Ain, D(sc.XB), ne, Np, &rf, I(sc.perm), I(sc.pinv), D(sc.P), I(sc.src),
                                              I(sc.piv), mask, st));
          TICC_RETURN_IF(gj_unpermute_masked(rf, D(sc.X), ne, Np, I(sc.pinv), mask, st));
        }
        *X = D(sc.X);
        *pv = I(sc.ipinv);
        return TICC_OK;
      };
      TICC_CHECK_HIP(hipStreamWaitEvent(st, ev_qm[qb], 0));
      for (int q = 0; q < nc; ++q) {
        const int s = s0 + q;
        const double h = hs_host[s];
        const double* Wm = Wm0 + s * wstride;
        const double* We = We0 + s * wstride;
        double* y1 = ys(sc.y1, s);
        double* y2 = ys(sc.y2, s);
        if (s == 0) {  // later sectors' A1 comes from the previous sector's k_form3
          TICC_CHECK_HIP(refvals(s));
          hipLaunchKernelGGL(k_form1, dim3(gelem, ne), dim3(256), 0, st, Yg, Ws0, Wm, h, y1, N, Np, A0, e0, per_block,
                             gstride);
          TICC_CHECK_LAUNCH();
          a_in[g] = A0;
          a_alt[g] = A1;
        }
        double* r1 = nullptr;
        const int* pv1 = nullptr;
        TICC_RETURN_IF(invert(a_in[g], a_alt[g], &r1, &pv1));
        // A2 goes where neither A1's inverse nor (mixed) A1 itself is needed again
        double* other = mixed ? a_alt[g] : ((r1 == A0) ? A1 : A0);
        hipLaunchKernelGGL(k_form2, dim3((Np + 3) / 4, ne), dim3(256), 0, st, r1, pv1, Qbuf + q * n2, y1, y2, N, Np, other,
                           e0, per_block, qstride);
        TICC_CHECK_LAUNCH();
        double* r2 = nullptr;
        const int* pv2 = nullptr;
        TICC_RETURN_IF(invert(other, mixed ? a_in[g] : r1, &r2, &pv2));
        NextSector nx{nullptr, nullptr, 0.0, y1, nullptr};
        if (s + 1 < nS) {  // form the next sector's A1 in a buffer X2 does not occupy
          TICC_CHECK_HIP(refvals(s + 1));
          double* nA = mixed ? a_in[g] : ((r2 == A0) ? A1 : A0);
          nx = NextSector{Ws0 + (s + 1) * wstride, Wm0 + (s + 1) * wstride, hs_host[s + 1], ys(sc.y1, s + 1), nA};
          if (!mixed) {
            a_in[g] = nA;
            a_alt[g] = r2;
          }
        }
        hipLaunchKernelGGL(k_form3, dim3(tn * (tn + 1) / 2, ne), dim3(256), 0, st, r2, pv2, We, Wm, h, y1, y2, N, Np,
                           half, Yg, e0, per_block, gstride, nx);
        TICC_CHECK_LAUNCH();
      }
      TICC_CHECK_HIP(hipEventRecord(ev_done[2 * g + qb], st));
    }
  }
  // join: the caller's stream waits for every group's last chunk
  const int last = (n_chunk - 1) & 1;
  for (int g = 0; g < G; ++g)
    if (g * per < nE) TICC_CHECK_HIP(hipStreamWaitEvent(stream, ev_done[2 * g + last], 0));
  TICC_RETURN_IF(unpad(Yp, Np, Y, (long long)N * N, N, nE, stream));
  if (as) return TICC_OK;  // no host synchronisation: the caller checks Y
  TICC_RETURN_IF(check_finite(Y, (long long)nE * N * N, flag, stream));
  int hflag = 0;
  unsigned long long hcount = 0;
  TICC_CHECK_HIP(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, stream));
  if (mixed) TICC_CHECK_HIP(hipMemcpyAsync(&hcount, fb_count, sizeof(hcount), hipMemcpyDeviceToHost, stream));
  TICC_CHECK_HIP(hipStreamSynchronize(stream));
  if (mixed) {
    t_mixed_inversions += 2LL * nS * nE;
    t_mixed_fallbacks += (long long)hcount;
  }
  if (hflag) return fail(TICC_ERR_NONFINITE, "LDMD sweep produced a non-finite log-derivative matrix");
  return TICC_OK;
}

}  // namespace ticc
