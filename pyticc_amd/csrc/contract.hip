// W(R) contraction and radial-matrix assembly (kernel A).
//
// For every radial point r and helicity block K with weighted basis rows
// G = [B_real | B_imag] (N_K x n_ang each):
//   V_K(r) = norm (B_real diag(V_r) B_real^T + B_imag diag(V_r) B_imag^T)
//   W(r)[a][b] = U[a][b] / R_r^2 + 2 mu V[a][b] (+ 2 mu E_int[a] on the diagonal)
// Cross-K elements of V are zero (W there is U/R^2 only).  One workgroup of
// CW waves computes one 32 x 32 tile (a-tile <= b-tile) with FP64 MFMA: each
// wave sums a contiguous CW-th of the angular grid, diag(V_r) applied to the A
// operand as it is loaded; the partial tiles meet in LDS and the first wave
// writes both mirror images, so W is exactly symmetric.  Splitting the
// angular sum keeps a small block's few tiles from running as a handful of
// long serial waves.
//
// With enough tiles x radial points to fill the chip, k_contract_lds instead
// gives each 4-wave workgroup one tile and 4 PW radial points: the tile's two
// row sets (32 rows x KC angular points each) and the points' V values are
// staged in LDS with coalesced row loads, once per angular chunk for all the
// points, and each wave runs the MFMAs of its PW points from there.  Per
// point the MFMA sequence (angular steps of 4 in order, the real rows and then
// the imaginary rows into one accumulator) is that of k_contract<1>, so the
// two give the same W bit for bit.
#include <cstdlib>
#include <cstring>

#include "ticc_common.h"

namespace ticc {
namespace {

__global__ void k_w_init(const double* __restrict__ U, const double* __restrict__ R, int N, int nR,
                         double* __restrict__ W) {
  const size_t n2 = (size_t)N * N;
  const size_t total = n2 * nR;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(idx / n2);
    const double RR = R[r];
    W[idx] = U[idx % n2] / (RR * RR);
  }
}

// tiles[t] = (row offset o, block size nk, a-tile, b-tile); CW waves per tile,
// each summing a CW-th of the angular grid
template <int CW>
__global__ __launch_bounds__(64 * CW) void k_contract(const int4* __restrict__ tiles, const int* __restrict__ chan,
                                                 const double* __restrict__ Br, const double* __restrict__ Bi,
                                                 int n_ang, double norm, const double* __restrict__ R,
                                                 const double* __restrict__ V, const double* __restrict__ U,
                                                 const double* __restrict__ Eint, double two_mu, int N,
                                                 double* __restrict__ W) {
  const int4 tl = tiles[blockIdx.x];
  const int r = blockIdx.y;
  const int o = tl.x, nk = tl.y;
  const int a0 = 32 * tl.z, b0 = 32 * tl.w;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lr = lane & 15;
  const int lq = lane >> 4;
  const double* Vr = V + (size_t)r * n_ang;
  // this wave's k-steps of 4 angular points: [ks0, ks1)
  const int n_steps = (n_ang + 3) / 4;
  const int ks0 = (wave * n_steps) / CW;
  const int ks1 = ((wave + 1) * n_steps) / CW;

  double4_t acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = double4_t{0.0, 0.0, 0.0, 0.0};

  const double* rows_a[2];
  const double* rows_b[2];
  bool va[2], vb[2];
  for (int part = 0; part < 2; ++part) {
    const double* Bm = (part == 0) ? Br : Bi;
    if (Bm == nullptr) break;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int la = a0 + 16 * q + lr;
      const int lb = b0 + 16 * q + lr;
      va[q] = la < nk;
      vb[q] = lb < nk;
      rows_a[q] = Bm + (size_t)(o + (va[q] ? la : 0)) * n_ang;
      rows_b[q] = Bm + (size_t)(o + (vb[q] ? lb : 0)) * n_ang;
    }
#pragma unroll(CW > 1 ? 4 : 1)
    for (int ks = ks0; ks < ks1; ++ks) {
      const int g = 4 * ks + lq;
      const bool vg = g < n_ang;
      const double vv = vg ? Vr[g] : 0.0;
      double a[2], b[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        a[q] = (va[q] && vg) ? rows_a[q][g] * vv : 0.0;
        b[q] = (vb[q] && vg) ? rows_b[q][g] : 0.0;
      }
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
          acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], b[bj], acc[bi][bj], 0, 0, 0);
    }
  }

  // partial tiles of waves 1 .. CW-1 -> wave 0
  __shared__ double red[CW > 1 ? CW - 1 : 1][16][64];
  if (CW > 1 && wave > 0) {
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int q = 0; q < 4; ++q) red[wave - 1][8 * bi + 4 * bj + q][lane] = acc[bi][bj][q];
  }
  if (CW > 1) __syncthreads();
  if (wave > 0) return;
#pragma unroll 1
  for (int w = 0; w < CW - 1; ++w)
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[bi][bj][q] += red[w][8 * bi + 4 * bj + q][lane];

  const double RR = R[r];
  const double R2 = RR * RR;
  double* Wr = W + (size_t)r * N * N;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int la = a0 + 16 * bi + mfma_drow(lq, q);
        const int lb = b0 + 16 * bj + lr;
        if (la >= nk || lb >= nk) continue;
        if (a0 == b0 && la > lb) continue;
        const int ca = chan[o + la];
        const int cb = chan[o + lb];
        double w = U[(size_t)ca * N + cb] / R2 + two_mu * (norm * acc[bi][bj][q]);
        if (ca == cb) w += two_mu * Eint[ca];
        Wr[(size_t)ca * N + cb] = w;
        Wr[(size_t)cb * N + ca] = w;
      }
}

// The contraction's form, chosen by size (0) or forced for comparisons by
// TICC_CONTRACT_FORM: "wave" (k_contract<1>), "lds1", "lds4" (k_contract_lds<1|4>),
// "split", "split8" or "split16" (k_contract<4|8|16>).
enum { CF_AUTO, CF_WAVE, CF_LDS1, CF_LDS4, CF_SPLIT, CF_SPLIT8, CF_SPLIT16 };
int contract_form() {
  const char* env = getenv("TICC_CONTRACT_FORM");
  if (!env) return CF_AUTO;
  if (!strcmp(env, "wave")) return CF_WAVE;
  if (!strcmp(env, "lds1")) return CF_LDS1;
  if (!strcmp(env, "lds4")) return CF_LDS4;
  if (!strcmp(env, "split")) return CF_SPLIT;
  if (!strcmp(env, "split8")) return CF_SPLIT8;
  if (!strcmp(env, "split16")) return CF_SPLIT16;
  return CF_AUTO;
}

constexpr int CKC = 32;  // angular points per LDS chunk

template <int PW>
__global__ __launch_bounds__(256, 2) void k_contract_lds(const int4* __restrict__ tiles, const int* __restrict__ chan,
                                                      const double* __restrict__ Br, const double* __restrict__ Bi,
                                                      int n_ang, double norm, int nR, const double* __restrict__ R,
                                                      const double* __restrict__ V, const double* __restrict__ U,
                                                      const double* __restrict__ Eint, double two_mu, int N,
                                                      double* __restrict__ W) {
  constexpr int P = 4 * PW;  // radial points per workgroup
  __shared__ double As[32][CKC + 1];
  __shared__ double Bs[32][CKC + 1];
  __shared__ double Vs[P][CKC];
  const int4 tl = tiles[blockIdx.x];
  const int o = tl.x, nk = tl.y;
  const int a0 = 32 * tl.z, b0 = 32 * tl.w;
  const int r0 = blockIdx.y * P;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int lr = lane & 15;
  const int lq = lane >> 4;
  // staging: each thread loads column (t & 31) of rows (t >> 5) + 8 i
  const int sc = t & 31;
  const int sr = t >> 5;

  double4_t acc[PW][2][2];
#pragma unroll
  for (int p = 0; p < PW; ++p)
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) acc[p][bi][bj] = double4_t{0.0, 0.0, 0.0, 0.0};

  for (int part = 0; part < 2; ++part) {
    const double* Bm = (part == 0) ? Br : Bi;
    if (Bm == nullptr) break;
    for (int k0 = 0; k0 < n_ang; k0 += CKC) {
      const int g = k0 + sc;
      const bool vg = g < n_ang;
      __syncthreads();  // the previous chunk has been consumed
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = sr + 8 * i;
        const int la = a0 + row, lb = b0 + row;
        As[row][sc] = (vg && la < nk) ? Bm[(size_t)(o + la) * n_ang + g] : 0.0;
        Bs[row][sc] = (vg && lb < nk) ? Bm[(size_t)(o + lb) * n_ang + g] : 0.0;
      }
      for (int idx = t; idx < P * CKC; idx += 256) {
        const int p = idx / CKC, c = idx % CKC;
        const int r = r0 + p;
        Vs[p][c] = (r < nR && k0 + c < n_ang) ? V[(size_t)r * n_ang + k0 + c] : 0.0;
      }
      __syncthreads();
#pragma unroll 2
      for (int ks = 0; ks < CKC / 4; ++ks) {
        const int c = 4 * ks + lq;
        double a[2], b[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          a[q] = As[16 * q + lr][c];
          b[q] = Bs[16 * q + lr][c];
        }
#pragma unroll
        for (int p = 0; p < PW; ++p) {
          const double vv = Vs[wave * PW + p][c];
#pragma unroll
          for (int bi = 0; bi < 2; ++bi) {
            const double av = a[bi] * vv;
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
              acc[p][bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b[bj], acc[p][bi][bj], 0, 0, 0);
          }
        }
      }
    }
  }

#pragma unroll
  for (int p = 0; p < PW; ++p) {
    const int r = r0 + wave * PW + p;
    if (r >= nR) continue;
    const double RR = R[r];
    const double R2 = RR * RR;
    double* Wr = W + (size_t)r * N * N;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int la = a0 + 16 * bi + mfma_drow(lq, q);
          const int lb = b0 + 16 * bj + lr;
          if (la >= nk || lb >= nk) continue;
          if (a0 == b0 && la > lb) continue;
          const int ca = chan[o + la];
          const int cb = chan[o + lb];
          double w = U[(size_t)ca * N + cb] / R2 + two_mu * (norm * acc[p][bi][bj][q]);
          if (ca == cb) w += two_mu * Eint[ca];
          Wr[(size_t)ca * N + cb] = w;
          Wr[(size_t)cb * N + ca] = w;
        }
  }
}

}  // namespace

int contract_w(int N, int n_ang, int n_block, const int* block_offsets, const int* block_channels,
               const double* Br, const double* Bi, double norm, int nR, const double* R, const double* V,
               const double* U, const double* Eint, double mu, double* W, hipStream_t stream) {
  // tile list (host) -> workspace (device)
  int n_tiles = 0;
  for (int k = 0; k < n_block; ++k) {
    const int nk = block_offsets[k + 1] - block_offsets[k];
    const int t = (nk + 31) / 32;
    n_tiles += t * (t + 1) / 2;
  }
  {
    long long total = (long long)nR * N * N;
    long long blocks = (total + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_w_init, dim3((int)blocks), dim3(256), 0, stream, U, R, N, nR, W);
    TICC_CHECK_LAUNCH();
  }
  if (n_tiles == 0 || nR == 0) return TICC_OK;
  int4* host_tiles = new int4[n_tiles];
  int p = 0;
  for (int k = 0; k < n_block; ++k) {
    const int o = block_offsets[k];
    const int nk = block_offsets[k + 1] - o;
    const int t = (nk + 31) / 32;
    for (int ta = 0; ta < t; ++ta)
      for (int tb = ta; tb < t; ++tb) host_tiles[p++] = make_int4(o, nk, ta, tb);
  }
  void* ws = nullptr;
  hipError_t e0 = hipStreamSynchronize(stream);  // earlier calls may still read the tile list
  if (e0 != hipSuccess) {
    delete[] host_tiles;
    return fail(TICC_ERR_HIP, "stream synchronize failed: %s", hipGetErrorString(e0));
  }
  int st = workspace(WS_CONTRACT, sizeof(int4) * n_tiles, &ws);
  if (st != TICC_OK) {
    delete[] host_tiles;
    return st;
  }
  hipError_t e = hipMemcpyAsync(ws, host_tiles, sizeof(int4) * n_tiles, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  delete[] host_tiles;
  if (e != hipSuccess) return fail(TICC_ERR_HIP, "tile upload failed: %s", hipGetErrorString(e));
  // The form depends on the number of radial points only, so that every block
  // contracted over the same window rounds its W the same way (a bucket of blocks
  // truncated from one basis takes the largest member's W for all of them, bit for
  // bit what each would contract alone).  Windows of 32 points or more: the
  // LDS-staged form, 16 points per workgroup (N = 800, 256 points: 5.5 ms against
  // 12.7 ms for one wave per tile and point).  Fewer points (short sweeps, a last
  // short window) leave too few workgroups for that, and each tile's angular sum is
  // split over 16 waves instead (17 points: 0.14-0.32 ms for 4-27 tiles, 1.5-2.2x
  // faster than four waves).
  int form = contract_form();
  if (form == CF_AUTO) form = nR >= 32 ? CF_LDS4 : CF_SPLIT16;
  if (form == CF_LDS4)
    hipLaunchKernelGGL(k_contract_lds<4>, dim3(n_tiles, (nR + 15) / 16), dim3(256), 0, stream,
                       static_cast<const int4*>(ws), block_channels, Br, Bi, n_ang, norm, nR, R, V, U, Eint,
                       2.0 * mu, N, W);
  else if (form == CF_LDS1)
    hipLaunchKernelGGL(k_contract_lds<1>, dim3(n_tiles, (nR + 3) / 4), dim3(256), 0, stream,
                       static_cast<const int4*>(ws), block_channels, Br, Bi, n_ang, norm, nR, R, V, U, Eint,
                       2.0 * mu, N, W);
  else if (form == CF_WAVE)
    hipLaunchKernelGGL(k_contract<1>, dim3(n_tiles, nR), dim3(64), 0, stream, static_cast<const int4*>(ws),
                       block_channels, Br, Bi, n_ang, norm, R, V, U, Eint, 2.0 * mu, N, W);
  else if (form == CF_SPLIT8)
    hipLaunchKernelGGL(k_contract<8>, dim3(n_tiles, nR), dim3(512), 0, stream, static_cast<const int4*>(ws),
                       block_channels, Br, Bi, n_ang, norm, R, V, U, Eint, 2.0 * mu, N, W);
  else if (form == CF_SPLIT16)
    hipLaunchKernelGGL(k_contract<16>, dim3(n_tiles, nR), dim3(1024), 0, stream, static_cast<const int4*>(ws),
                       block_channels, Br, Bi, n_ang, norm, R, V, U, Eint, 2.0 * mu, N, W);
  else
    hipLaunchKernelGGL(k_contract<4>, dim3(n_tiles, nR), dim3(256), 0, stream, static_cast<const int4*>(ws),
                       block_channels, Br, Bi, n_ang, norm, R, V, U, Eint, 2.0 * mu, N, W);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

}  // namespace ticc
