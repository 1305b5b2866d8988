// Batched blocked Gauss-Jordan inversion with partial pivoting (gfx950).
//
// A batch of padded Np x Np FP64 matrices is inverted in Np / NB block steps.
// Step k0 (columns [k0, k0+NB)):
//   1. k_gj_panel   (one workgroup per matrix): in-place Gauss-Jordan on the
//      Np x NB column panel, pivot rows chosen by partial pivoting among the
//      rows not yet used (LAPACK idamax order: largest |a|, first position on
//      ties).  The panel lives in registers, up to 2 rows per thread.  Outputs the
//      transformed panel in output-row order, the row source map of the step
//      and the NB pivot rows.
//   2. k_gj_update  (tiles x matrices): the rank-NB update
//        Out[i][:] = Cin_i + Pout[i][:] . X[piv][:],  Cin_i = X[src(i)] (0 on
//      the pivot block), panel columns replaced by Pout -- an FP64 MFMA GEMM
//      with K = NB, reading X and writing the ping-pong partner buffer.
// Row swaps compose into perm (original row at each position); the inverse of
// the matrix is Out with its columns gathered through pinv = perm^-1, which the
// consumers fuse into their own reads.
#include <mutex>
#include <vector>

#include "ticc_common.h"

namespace ticc {

namespace {

constexpr int NB = GJ_NB;

struct PivotCand {
  double v;
  int pos;
  int row;
};

__device__ __forceinline__ bool better(double v1, int p1, double v2, int p2) {
  // larger magnitude wins; equal magnitudes: smaller position (idamax order).
  return (v1 > v2) || (v1 == v2 && p1 < p2);
}

__device__ __forceinline__ PivotCand wave_argmax(PivotCand c) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    double ov = __shfl_xor(c.v, off, 64);
    int op = __shfl_xor(c.pos, off, 64);
    int orow = __shfl_xor(c.row, off, 64);
    if (better(ov, op, c.v, c.pos)) {
      c.v = ov;
      c.pos = op;
      c.row = orow;
    }
  }
  return c;
}

// Wave-wide argmax of (value, position) -- larger value wins, smaller position on
// ties -- without LDS round trips: rotations within each 16-lane row by DPP
// (row_ror 8, 4, 2, 1 leave the row maximum in every lane of the row), then the
// four row results are read with v_readlane and combined in uniform registers.
template <int CTRL>
__device__ __forceinline__ void dpp_argmax_step(double& v, int& p) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), CTRL, 0xF, 0xF, false);
  const int op = __builtin_amdgcn_update_dpp(0, p, CTRL, 0xF, 0xF, false);
  const double ov = __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  if (better(ov, op, v, p)) {
    v = ov;
    p = op;
  }
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ void wave_argmax_dpp(double& v, int& p) {
  dpp_argmax_step<0x128>(v, p);  // row_ror:8
  dpp_argmax_step<0x124>(v, p);  // row_ror:4
  dpp_argmax_step<0x122>(v, p);  // row_ror:2
  dpp_argmax_step<0x121>(v, p);  // row_ror:1
  double bv = readlane_f64(v, 0);
  int bp = __builtin_amdgcn_readlane(p, 0);
#pragma unroll
  for (int r = 1; r < 4; ++r) {
    const double rv = readlane_f64(v, 16 * r);
    const int rp = __builtin_amdgcn_readlane(p, 16 * r);
    if (better(rv, rp, bv, bp)) {
      bv = rv;
      bp = rp;
    }
  }
  v = bv;
  p = bp;
}

// Panel: one workgroup per matrix; thread t owns rows t + THREADS q
// (q < RPT, Np <= THREADS RPT) with their NB panel values in registers and
// its rows' current positions.  Per column: each wave reduces its best
// candidate (largest |a|, first position on ties) and its winning lane
// publishes the candidate row's values in a double-buffered LDS slot; after
// the column's single barrier every thread picks the global winner, applies
// the row swap to its own position records and performs the in-place
// Gauss-Jordan update, streaming the pivot row from LDS.
template <int THREADS, int RPT>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(THREADS / 256, THREADS / 256))) void k_gj_panel(const double* __restrict__ X, long long bstride, int ld, int col0, int Np, int k0,
                                                      double* __restrict__ Pout, int* __restrict__ src,
                                                      int* __restrict__ piv, int* __restrict__ perm,
                                                      int* __restrict__ pinv) {
  constexpr int NW = THREADS / 64;
  __shared__ double s_vals[2][NW][NB];
  __shared__ double s_v[2][NW];
  __shared__ int s_pos[2][NW];
  __shared__ int s_row[2][NW];

  const int b = blockIdx.x;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const double* Xb = X + (size_t)b * bstride;

  double v[RPT][NB];
  int pos[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = t + THREADS * q;
    if (r < Np) {
      const double2* rp = reinterpret_cast<const double2*>(Xb + (size_t)r * ld + col0);
#pragma unroll
      for (int l = 0; l < NB / 2; ++l) {
        const double2 d = rp[l];
        v[q][2 * l] = d.x;
        v[q][2 * l + 1] = d.y;
      }
      pos[q] = r;
    } else {
#pragma unroll
      for (int l = 0; l < NB; ++l) v[q][l] = 0.0;
      pos[q] = -1;
    }
  }

#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int c = k0 + j;
    const int buf = j & 1;
    // candidate: largest |a| among rows at positions >= c (NaN ranks as +inf so
    // a column always has a pivot and the NaN propagates to the result)
    double cv = -1.0;
    int cp = 0x7fffffff;
    int cq = -1;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      if (pos[q] >= c) {
        double a = fabs(v[q][j]);
        if (a != a) a = __builtin_inf();
        if (better(a, pos[q], cv, cp)) {
          cv = a;
          cp = pos[q];
          cq = q;
        }
      }
    }
    double wv = cv;
    int wp = cp;
    wave_argmax_dpp(wv, wp);
    if (cq >= 0 && cp == wp) {  // this lane holds the wave's winning row
#pragma unroll
      for (int q = 0; q < RPT; ++q)
        if (q == cq) {
#pragma unroll
          for (int l = 0; l < NB; ++l) s_vals[buf][wave][l] = v[q][l];
          s_row[buf][wave] = t + THREADS * q;
        }
    }
    if (lane == 0) {
      s_v[buf][wave] = wv;
      s_pos[buf][wave] = wp;
    }
    __syncthreads();
    double best_v = s_v[buf][0];
    int bp = s_pos[buf][0];
    int bw = 0;
#pragma unroll
    for (int w = 1; w < NW; ++w) {
      if (better(s_v[buf][w], s_pos[buf][w], best_v, bp)) {
        best_v = s_v[buf][w];
        bp = s_pos[buf][w];
        bw = w;
      }
    }
    const int pr = s_row[buf][bw];
    const double* prow = s_vals[buf][bw];
    const double inv = 1.0 / prow[j];
    // non-pivot rows: a_rl -= (a_rj / p) p_l, a_rj = -a_rj / p; the pivot row is
    // rescaled afterwards in a branch taken by a single lane.
    double g[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      if (pos[q] == c) pos[q] = bp;
      g[q] = v[q][j] * inv;
    }
#pragma unroll
    for (int l = 0; l < NB; ++l) {
      const double pl = prow[l];
#pragma unroll
      for (int q = 0; q < RPT; ++q) v[q][l] = fma(-g[q], pl, v[q][l]);
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      v[q][j] = -g[q];
      if (t + THREADS * q == pr) {
        pos[q] = c;
#pragma unroll
        for (int l = 0; l < NB; ++l) v[q][l] = prow[l] * inv;
        v[q][j] = inv;
      }
    }
  }

  // Outputs in output-position order.
  int pold[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = t + THREADS * q;
    pold[q] = 0;
    if (pos[q] < 0) continue;
    const int p = pos[q];
    double2* op = reinterpret_cast<double2*>(Pout + ((size_t)b * Np + p) * NB);
#pragma unroll
    for (int l = 0; l < NB / 2; ++l) op[l] = make_double2(v[q][2 * l], v[q][2 * l + 1]);
    src[(size_t)b * Np + p] = r;
    if (p >= k0 && p < k0 + NB) piv[b * NB + (p - k0)] = r;
    pold[q] = (k0 == 0) ? r : perm[(size_t)b * Np + r];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    if (pos[q] < 0) continue;
    perm[(size_t)b * Np + pos[q]] = pold[q];
    if (k0 + NB >= Np) pinv[(size_t)b * Np + pold[q]] = pos[q];
  }
}

// Update: 256 threads = 4 waves on a 64 x 64 output tile, each wave 32 x 32
// (2 x 2 FP64 MFMA 16x16x4 blocks, K = NB).  The Pout rows and the NB pivot
// rows of the tile are staged through LDS with coalesced loads; the gathered
// input rows X[src(i)] are loaded straight into the accumulators before the
// barrier.  Panel columns are overwritten with the transformed panel.
__global__ __launch_bounds__(256) void k_gj_update(const double* __restrict__ X, double* __restrict__ Out, int Np,
                                                   int k0, const double* __restrict__ Pout,
                                                   const int* __restrict__ src, const int* __restrict__ piv) {
  __shared__ double sA[64][NB + 1];
  __shared__ double sB[NB][64];
  const int tiles = (Np + 63) / 64;
  const int I0 = (blockIdx.x / tiles) * 64;
  const int J0 = (blockIdx.x % tiles) * 64;
  const int b = blockIdx.y;
  const int t = threadIdx.x;
  const int wave = t >> 6;
  const int lane = t & 63;
  const int i0 = I0 + (wave >> 1) * 32;
  const int j0 = J0 + (wave & 1) * 32;
  const bool active = (i0 < Np) && (j0 < Np);
  const bool panel = (j0 == k0);

  const double* Xb = X + (size_t)b * Np * Np;
  double* Ob = Out + (size_t)b * Np * Np;
  const double* Pb = Pout + (size_t)b * Np * NB;
  const int* srcb = src + (size_t)b * Np;
  const int* pivb = piv + b * NB;
  const int lr = lane & 15;
  const int lq = lane >> 4;

  // stage A = Pout[I0:I0+64][0:NB] and B = X[piv][J0:J0+64]
  {
    const int row = t >> 2, part = t & 3;
    const int gi = I0 + row;
    double2 a[4];
#pragma unroll
    for (int h = 0; h < 4; ++h) a[h] = (gi < Np) ? reinterpret_cast<const double2*>(Pb + (size_t)gi * NB + part * 8)[h] : make_double2(0.0, 0.0);
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      sA[row][part * 8 + 2 * h] = a[h].x;
      sA[row][part * 8 + 2 * h + 1] = a[h].y;
    }
    const int k = t >> 3, seg = t & 7;
    const int gj = J0 + seg * 8;
    const double* brow = Xb + (size_t)pivb[k] * Np + gj;
    double2 bb[4];
#pragma unroll
    for (int h = 0; h < 4; ++h) bb[h] = (gj < Np) ? reinterpret_cast<const double2*>(brow)[h] : make_double2(0.0, 0.0);
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      sB[k][seg * 8 + 2 * h] = bb[h].x;
      sB[k][seg * 8 + 2 * h + 1] = bb[h].y;
    }
  }

  double4_t acc[2][2];
  if (active && !panel) {
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + 16 * bi + mfma_drow(lq, r);
        const bool pivblk = (row >= k0) && (row < k0 + NB);
        const double* xr = Xb + (size_t)srcb[row] * Np + j0 + lr;
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj][r] = pivblk ? 0.0 : xr[16 * bj];
      }
  }
  __syncthreads();
  if (!active) return;
  const int li0 = i0 - I0;
  const int lj0 = j0 - J0;
  if (panel) {
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          const int lrow = li0 + 16 * bi + mfma_drow(lq, r);
          const int col = 16 * bj + lr;
          Ob[(size_t)(I0 + lrow) * Np + k0 + col] = sA[lrow][col];
        }
    return;
  }
#pragma unroll
  for (int kk = 0; kk < NB / 4; ++kk) {
    const int kq = 4 * kk + lq;
    double a[2], bv[2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) a[bi] = sA[li0 + 16 * bi + lr][kq];
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) bv[bj] = sB[kq][lj0 + 16 * bj + lr];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
        acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], bv[bj], acc[bi][bj], 0, 0, 0);
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = i0 + 16 * bi + mfma_drow(lq, r);
      double* orow = Ob + (size_t)row * Np + j0 + lr;
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) orow[16 * bj] = acc[bi][bj][r];
    }
}

// Small panel GEMM (FP64 MFMA) used by the paired block step:
//   Out[i][0:NB] = mask(i) C[gC(i)][colC : colC+NB] + A[i][0:NB] . B[gB(r)][colB : colB+NB],
//   mask(i) = 0 for i in [mlo, mlo+NB), 1 otherwise.
// One workgroup (4 waves) per 64 rows of one matrix, each wave 16 rows x NB.
__global__ __launch_bounds__(256) void k_panel_gemm(const double* __restrict__ A, const double* __restrict__ C,
                                                    long long cb, int ldC, int colC, const int* __restrict__ gC,
                                                    const double* __restrict__ Bm, long long bb, int ldB, int colB,
                                                    const int* __restrict__ gB, int mlo, int Np,
                                                    double* __restrict__ Out) {
  __shared__ double sB[NB][NB + 1];
  const int b = blockIdx.y;
  const int t = threadIdx.x;
  const int wave = t >> 6, lane = t & 63, lr = lane & 15, lq = lane >> 4;
  const double* Ab = A + (size_t)b * Np * NB;
  const double* Cb = C + (size_t)b * cb;
  const double* Bb = Bm + (size_t)b * bb;
  const int* gCb = gC + (size_t)b * Np;
  const int* gBb = gB + b * NB;
  {  // stage B: NB gathered rows x NB columns
    const int r = t >> 3, seg = t & 7;  // 32 rows x 8 segments of 4
    const double* br = Bb + (size_t)gBb[r] * ldB + colB + seg * 4;
#pragma unroll
    for (int h = 0; h < 4; ++h) sB[r][seg * 4 + h] = br[h];
  }
  const int row0 = blockIdx.x * 64 + wave * 16;
  double4_t acc[2];
  if (row0 < Np) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = row0 + mfma_drow(lq, r);
      const bool m = (row >= mlo) && (row < mlo + NB);
      const double* cr = Cb + (size_t)gCb[row] * ldC + colC + lr;
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) acc[bj][r] = m ? 0.0 : cr[16 * bj];
    }
  }
  __syncthreads();
  if (row0 >= Np) return;
#pragma unroll
  for (int kk = 0; kk < NB / 4; ++kk) {
    const int kq = 4 * kk + lq;
    const double a = Ab[(size_t)(row0 + lr) * NB + kq];
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
      acc[bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sB[kq][16 * bj + lr], acc[bj], 0, 0, 0);
  }
  double* Ob = Out + (size_t)b * Np * NB;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = row0 + mfma_drow(lq, r);
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) Ob[(size_t)row * NB + 16 * bj + lr] = acc[bj][r];
  }
}

// Paired block step k, k+1 (columns [k0, k0+NB) and [k1, k1+NB), k1 = k0+NB):
//   X_{k+2}[i][c] = mask(i) X_k[s1(s2(i))][c]
//                   + A1[i] . X_k[p1][c] + P2[i] . X_k[s1(p2)][c]
// with mask(i) = 0 if i is in the step-(k+1) pivot block or s2(i) in the step-k
// pivot block; panel columns are replaced by A1 (step k) and P2 (step k+1).
// Rank-2NB FP64 MFMA update: one pass over X for two Gauss-Jordan steps.
__global__ __launch_bounds__(256) void k_gj_update2(const double* __restrict__ X, double* __restrict__ Out, int Np,
                                                    int k0, const double* __restrict__ A1,
                                                    const double* __restrict__ P2, const int* __restrict__ s1,
                                                    const int* __restrict__ s2, const int* __restrict__ p1,
                                                    const int* __restrict__ p2) {
  constexpr int K2 = 2 * NB;
  __shared__ double sB[K2][64];
  const int tiles = (Np + 63) / 64;
  const int I0 = (blockIdx.x / tiles) * 64;
  const int J0 = (blockIdx.x % tiles) * 64;
  const int b = blockIdx.y;
  const int t = threadIdx.x;
  const int wave = t >> 6, lane = t & 63, lr = lane & 15, lq = lane >> 4;
  const int i0 = I0 + (wave >> 1) * 32;
  const int j0 = J0 + (wave & 1) * 32;
  const int k1 = k0 + NB;
  const bool active = (i0 < Np) && (j0 < Np);
  const double* Xb = X + (size_t)b * Np * Np;
  double* Ob = Out + (size_t)b * Np * Np;
  const double* A1b = A1 + (size_t)b * Np * NB;
  const double* P2b = P2 + (size_t)b * Np * NB;
  const int* s1b = s1 + (size_t)b * Np;
  const int* s2b = s2 + (size_t)b * Np;

  {  // stage the 2NB gathered B rows of the tile's 64 columns
    const int k = t >> 2, part = t & 3;  // 64 rows x 4 parts of 16 columns
    const int src_row = (k < NB) ? p1[b * NB + k] : s1b[p2[b * NB + (k - NB)]];
    const int gj = J0 + part * 16;
    const double2* br = reinterpret_cast<const double2*>(Xb + (size_t)src_row * Np + gj);
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      const double2 d = (gj < Np) ? br[h] : make_double2(0.0, 0.0);
      sB[k][part * 16 + 2 * h] = d.x;
      sB[k][part * 16 + 2 * h + 1] = d.y;
    }
  }
  const bool pan1 = (j0 == k0), pan2 = (j0 == k1);
  double4_t acc[2][2];
  if (active && !pan1 && !pan2) {
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + 16 * bi + mfma_drow(lq, r);
        const int m = s2b[row];
        const bool zero = (row >= k1 && row < k1 + NB) || (m >= k0 && m < k0 + NB);
        const double* xr = Xb + (size_t)s1b[m] * Np + j0 + lr;
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj][r] = zero ? 0.0 : xr[16 * bj];
      }
  }
  __syncthreads();
  if (!active) return;
  if (pan1 || pan2) {
    const double* Pb = pan1 ? A1b : P2b;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + 16 * bi + mfma_drow(lq, r);
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) Ob[(size_t)row * Np + j0 + 16 * bj + lr] = Pb[(size_t)row * NB + 16 * bj + lr];
      }
    return;
  }
  const int lj0 = j0 - J0;
#pragma unroll
  for (int kk = 0; kk < K2 / 4; ++kk) {
    const int kq = 4 * kk + lq;
    const double* Asrc = (kk < NB / 4) ? A1b : P2b;
    const int kc = (kk < NB / 4) ? kq : kq - NB;
    double a[2], bv[2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) a[bi] = Asrc[(size_t)(i0 + 16 * bi + lr) * NB + kc];
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) bv[bj] = sB[kq][lj0 + 16 * bj + lr];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
        acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], bv[bj], acc[bi][bj], 0, 0, 0);
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = i0 + 16 * bi + mfma_drow(lq, r);
      double* orow = Ob + (size_t)row * Np + j0 + lr;
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) orow[16 * bj] = acc[bi][bj][r];
    }
}

__global__ void k_unpermute(const double* __restrict__ in, double* __restrict__ out, int Np,
                            const int* __restrict__ pinv) {
  const int b = blockIdx.y;
  const size_t n2 = (size_t)Np * Np;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n2; idx += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(idx / Np);
    const int c = (int)(idx % Np);
    out[b * n2 + idx] = in[b * n2 + (size_t)r * Np + pinv[(size_t)b * Np + c]];
  }
}

// Live timing of the rank-NB update kernel with HIP events: every `every`-th
// inversion has each of its update launches bracketed by an event pair.
struct UpdateProfile {
  std::mutex mu;
  bool on = false;
  int every = 1;
  long long calls = 0;
  std::vector<hipEvent_t> events;  // pairs: [2i] start, [2i+1] stop
  std::vector<int> batch;          // matrices in each timed launch
  size_t used = 0;
};
UpdateProfile g_prof;

}  // namespace

int profile_begin(int every, int max_pairs) {
  std::lock_guard<std::mutex> lock(g_prof.mu);
  g_prof.every = every < 1 ? 1 : every;
  g_prof.calls = 0;
  g_prof.used = 0;
  while (g_prof.events.size() < (size_t)2 * max_pairs) {
    hipEvent_t e;
    TICC_CHECK_HIP(hipEventCreate(&e));
    g_prof.events.push_back(e);
  }
  g_prof.batch.resize(g_prof.events.size() / 2);
  g_prof.on = true;
  return TICC_OK;
}

int profile_end(double* total_ms, int* count, long long* total_batch) {
  std::lock_guard<std::mutex> lock(g_prof.mu);
  g_prof.on = false;
  TICC_CHECK_HIP(hipDeviceSynchronize());
  double tot = 0.0;
  long long nb = 0;
  for (size_t i = 0; i < g_prof.used; ++i) {
    nb += g_prof.batch[i];
    float ms = 0.0f;
    TICC_CHECK_HIP(hipEventElapsedTime(&ms, g_prof.events[2 * i], g_prof.events[2 * i + 1]));
    tot += ms;
  }
  *total_ms = tot;
  *count = (int)g_prof.used;
  *total_batch = nb;
  g_prof.used = 0;
  return TICC_OK;
}

namespace {
int launch_panel(const double* X, long long bstride, int ld, int col0, int batch, int Np, int k0, double* pout,
                 int* src, int* piv, int* perm, int* pinv, hipStream_t stream) {
  if (Np <= 256)
    hipLaunchKernelGGL((k_gj_panel<256, 1>), dim3(batch), dim3(256), 0, stream, X, bstride, ld, col0, Np, k0, pout, src,
                       piv, perm, pinv);
  else if (Np <= 512)
    hipLaunchKernelGGL((k_gj_panel<512, 1>), dim3(batch), dim3(512), 0, stream, X, bstride, ld, col0, Np, k0, pout, src,
                       piv, perm, pinv);
  else
    hipLaunchKernelGGL((k_gj_panel<512, 2>), dim3(batch), dim3(512), 0, stream, X, bstride, ld, col0, Np, k0, pout, src,
                       piv, perm, pinv);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}
}  // namespace

int gj_invert_raw(double* bufA, double* bufB, int batch, int Np, double** result, int* perm, int* pinv,
                  double* pout, int* src, int* piv, hipStream_t stream) {
  if (Np % NB != 0 || Np > MAX_NP || Np <= 0 || batch <= 0)
    return fail(TICC_ERR_VALUE, "gj_invert_raw: invalid Np=%d batch=%d", Np, batch);
  // scratch carve-up (sizes per GJ_POUT_DOUBLES / GJ_SRC_INTS / GJ_PIV_INTS)
  const size_t pstride = (size_t)batch * Np * NB;
  double* P1 = pout;
  double* P2 = pout + pstride;
  double* Z = pout + 2 * pstride;
  double* A1 = pout + 3 * pstride;
  int* s1 = src;
  int* s2 = src + (size_t)batch * Np;
  int* p1 = piv;
  int* p2 = piv + batch * NB;
  const long long n2 = (long long)Np * Np;
  double* cur = bufA;
  double* nxt = bufB;
  const int tiles = (Np + 63) / 64;
  const int rtiles = (Np + 63) / 64;
  bool sample = false;
  if (g_prof.on) {
    std::lock_guard<std::mutex> lock(g_prof.mu);
    sample = g_prof.on && (g_prof.calls++ % g_prof.every == 0) &&
             (g_prof.used + (size_t)(Np / NB) <= g_prof.events.size() / 2);
  }
  const int steps = Np / NB;
  for (int k = 0; k < steps;) {
    const int k0 = k * NB;
    TICC_RETURN_IF(launch_panel(cur, n2, Np, k0, batch, Np, k0, P1, s1, p1, perm, pinv, stream));
    size_t slot = 0;
    if (k + 1 == steps) {  // trailing single step
      if (sample) {
        slot = g_prof.used++;
        g_prof.batch[slot] = batch;
        TICC_CHECK_HIP(hipEventRecord(g_prof.events[2 * slot], stream));
      }
      hipLaunchKernelGGL(k_gj_update, dim3(tiles * tiles, batch), dim3(256), 0, stream, cur, nxt, Np, k0, P1, s1, p1);
      TICC_CHECK_LAUNCH();
      if (sample) TICC_CHECK_HIP(hipEventRecord(g_prof.events[2 * slot + 1], stream));
      k += 1;
    } else {
      const int k1 = k0 + NB;
      // look-ahead panel columns of step k+1: Z = X_{k+1}[:, k1:k1+NB]
      hipLaunchKernelGGL(k_panel_gemm, dim3(rtiles, batch), dim3(256), 0, stream, P1, cur, n2, Np, k1, s1, cur, n2, Np,
                         k1, p1, k0, Np, Z);
      TICC_CHECK_LAUNCH();
      TICC_RETURN_IF(launch_panel(Z, (long long)Np * NB, NB, 0, batch, Np, k1, P2, s2, p2, perm, pinv, stream));
      // combined step-k coefficients A1 = mask P1[s2] + P2 . P1[p2]
      hipLaunchKernelGGL(k_panel_gemm, dim3(rtiles, batch), dim3(256), 0, stream, P2, P1, (long long)Np * NB, NB, 0, s2,
                         P1, (long long)Np * NB, NB, 0, p2, k1, Np, A1);
      TICC_CHECK_LAUNCH();
      if (sample) {
        slot = g_prof.used++;
        g_prof.batch[slot] = 2 * batch;  // one launch performs two block steps
        TICC_CHECK_HIP(hipEventRecord(g_prof.events[2 * slot], stream));
      }
      hipLaunchKernelGGL(k_gj_update2, dim3(tiles * tiles, batch), dim3(256), 0, stream, cur, nxt, Np, k0, A1, P2, s1,
                         s2, p1, p2);
      TICC_CHECK_LAUNCH();
      if (sample) TICC_CHECK_HIP(hipEventRecord(g_prof.events[2 * slot + 1], stream));
      k += 2;
    }
    double* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  *result = cur;
  return TICC_OK;
}

int gj_unpermute(const double* in, double* out, int batch, int Np, const int* pinv, hipStream_t stream) {
  const long long n2 = (long long)Np * Np;
  int blocks = (int)((n2 + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_unpermute, dim3(blocks, batch), dim3(256), 0, stream, in, out, Np, pinv);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

}  // namespace ticc
