// Batched blocked Gauss-Jordan inversion with partial pivoting (gfx950).
//
// A batch of padded Np x Np FP64 matrices is inverted in Np / NB block steps.
// Step k0 (columns [k0, k0+NB)):
//   1. k_gj_panel   (one workgroup per matrix): in-place Gauss-Jordan on the
//      Np x NB column panel, pivot rows chosen by partial pivoting among the
//      rows not yet used (LAPACK idamax order: largest |a|, first position on
//      ties).  The panel lives in registers, one row per thread.  Outputs the
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

// One workgroup per matrix; thread t owns row t (Np <= THREADS).
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_gj_panel(const double* __restrict__ X, int Np, int k0,
                                                       double* __restrict__ Pout, int* __restrict__ src,
                                                       int* __restrict__ piv, int* __restrict__ perm,
                                                       int* __restrict__ pinv) {
  constexpr int NWAVE = THREADS / 64;
  __shared__ int pos2row[THREADS];
  __shared__ int row2pos[THREADS];
  __shared__ double prow[NB];
  __shared__ double red_v[NWAVE];
  __shared__ int red_p[NWAVE];
  __shared__ int red_r[NWAVE];
  __shared__ int s_pr;

  const int b = blockIdx.x;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const double* Xb = X + (size_t)b * Np * Np;
  const bool own = t < Np;

  double v[NB];
  if (own) {
    const double2* rp = reinterpret_cast<const double2*>(Xb + (size_t)t * Np + k0);
#pragma unroll
    for (int l = 0; l < NB / 2; ++l) {
      double2 q = rp[l];
      v[2 * l] = q.x;
      v[2 * l + 1] = q.y;
    }
  } else {
#pragma unroll
    for (int l = 0; l < NB; ++l) v[l] = 0.0;
  }
  if (t < THREADS) {
    pos2row[t] = t;
    row2pos[t] = t;
  }
  __syncthreads();

#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int c = k0 + j;
    PivotCand cand{-1.0, 0x7fffffff, -1};
    if (own) {
      const int p = row2pos[t];
      if (p >= c) {
        cand.v = fabs(v[j]);
        cand.pos = p;
        cand.row = t;
      }
    }
    cand = wave_argmax(cand);
    if (lane == 0) {
      red_v[wave] = cand.v;
      red_p[wave] = cand.pos;
      red_r[wave] = cand.row;
    }
    __syncthreads();
    if (t == 0) {
      PivotCand best{red_v[0], red_p[0], red_r[0]};
      for (int w = 1; w < NWAVE; ++w)
        if (better(red_v[w], red_p[w], best.v, best.pos)) best = PivotCand{red_v[w], red_p[w], red_r[w]};
      if (best.row < 0) {  // no finite candidate (NaN column): keep the current row
        best.row = pos2row[c];
        best.pos = c;
      }
      const int pr = best.row;
      const int bp = best.pos;
      const int rc = pos2row[c];
      pos2row[c] = pr;
      pos2row[bp] = rc;
      row2pos[pr] = c;
      row2pos[rc] = bp;
      s_pr = pr;
    }
    __syncthreads();
    const int pr = s_pr;
    if (t == pr) {
#pragma unroll
      for (int l = 0; l < NB; ++l) prow[l] = v[l];
    }
    __syncthreads();
    if (own) {
      const double inv = 1.0 / prow[j];
      if (t == pr) {
#pragma unroll
        for (int l = 0; l < NB; ++l) v[l] = (l == j) ? inv : prow[l] * inv;
      } else {
        const double f = v[j];
#pragma unroll
        for (int l = 0; l < NB; ++l) v[l] = (l == j) ? -f * inv : v[l] - f * (prow[l] * inv);
      }
    }
  }
  __syncthreads();

  // Outputs in output-position order.
  if (own) {
    const int p = row2pos[t];
    double2* op = reinterpret_cast<double2*>(Pout + ((size_t)b * Np + p) * NB);
#pragma unroll
    for (int l = 0; l < NB / 2; ++l) op[l] = make_double2(v[2 * l], v[2 * l + 1]);
    src[(size_t)b * Np + t] = pos2row[t];
  }
  if (t < NB) piv[b * NB + t] = pos2row[k0 + t];
  int pnew = 0;
  if (own) pnew = (k0 == 0) ? pos2row[t] : perm[(size_t)b * Np + pos2row[t]];
  __syncthreads();
  if (own) {
    perm[(size_t)b * Np + t] = pnew;
    if (k0 + NB >= Np) pinv[(size_t)b * Np + pnew] = t;
  }
}

// Update: 256 threads = 4 waves, each wave one 32 x 32 output tile of a 64 x 64
// workgroup tile; FP64 MFMA 16x16x4 (2 x 2 blocks per wave, K = NB).
// MFMA operand layout: see mfma_drow in ticc_common.h.
__global__ __launch_bounds__(256) void k_gj_update(const double* __restrict__ X, double* __restrict__ Out, int Np,
                                                   int k0, const double* __restrict__ Pout,
                                                   const int* __restrict__ src, const int* __restrict__ piv) {
  const int tiles = (Np + 63) / 64;
  const int tile = blockIdx.x;
  const int b = blockIdx.y;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int i0 = (tile / tiles) * 64 + (wave >> 1) * 32;
  const int j0 = (tile % tiles) * 64 + (wave & 1) * 32;
  if (i0 >= Np || j0 >= Np) return;

  const double* Xb = X + (size_t)b * Np * Np;
  double* Ob = Out + (size_t)b * Np * Np;
  const double* Pb = Pout + (size_t)b * Np * NB;
  const int* srcb = src + (size_t)b * Np;
  const int* pivb = piv + b * NB;
  const int lr = lane & 15;
  const int lq = lane >> 4;

  if (j0 == k0) {
    // panel columns: copy the transformed panel.
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          const int row = i0 + 16 * bi + mfma_drow(lq, r);
          const int col = 16 * bj + lr;
          Ob[(size_t)row * Np + k0 + col] = Pb[(size_t)row * NB + col];
        }
    return;
  }

  double4_t acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + 16 * bi + mfma_drow(lq, r);
        const int col = j0 + 16 * bj + lr;
        const bool pivblk = (row >= k0) && (row < k0 + NB);
        acc[bi][bj][r] = pivblk ? 0.0 : Xb[(size_t)srcb[row] * Np + col];
      }

#pragma unroll
  for (int kk = 0; kk < NB / 4; ++kk) {
    const int kcol = 4 * kk + lq;
    const int prow_idx = pivb[kcol];
    double a[2], bv[2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) a[bi] = Pb[(size_t)(i0 + 16 * bi + lr) * NB + kcol];
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) bv[bj] = Xb[(size_t)prow_idx * Np + j0 + 16 * bj + lr];
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
        Ob[(size_t)row * Np + col] = acc[bi][bj][r];
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
  g_prof.on = true;
  return TICC_OK;
}

int profile_end(double* total_ms, int* count) {
  std::lock_guard<std::mutex> lock(g_prof.mu);
  g_prof.on = false;
  TICC_CHECK_HIP(hipDeviceSynchronize());
  double tot = 0.0;
  for (size_t i = 0; i < g_prof.used; ++i) {
    float ms = 0.0f;
    TICC_CHECK_HIP(hipEventElapsedTime(&ms, g_prof.events[2 * i], g_prof.events[2 * i + 1]));
    tot += ms;
  }
  *total_ms = tot;
  *count = (int)g_prof.used;
  g_prof.used = 0;
  return TICC_OK;
}

int gj_invert_raw(double* bufA, double* bufB, int batch, int Np, double** result, int* perm, int* pinv,
                  double* pout, int* src, int* piv, hipStream_t stream) {
  if (Np % NB != 0 || Np > MAX_NP || Np <= 0 || batch <= 0)
    return fail(TICC_ERR_VALUE, "gj_invert_raw: invalid Np=%d batch=%d", Np, batch);
  double* cur = bufA;
  double* nxt = bufB;
  const int tiles = (Np + 63) / 64;
  bool sample = false;
  if (g_prof.on) {
    std::lock_guard<std::mutex> lock(g_prof.mu);
    sample = g_prof.on && (g_prof.calls++ % g_prof.every == 0) &&
             (g_prof.used + (size_t)(Np / NB) <= g_prof.events.size() / 2);
  }
  for (int k0 = 0; k0 < Np; k0 += NB) {
    if (Np <= 256)
      hipLaunchKernelGGL(k_gj_panel<256>, dim3(batch), dim3(256), 0, stream, cur, Np, k0, pout, src, piv, perm, pinv);
    else
      hipLaunchKernelGGL(k_gj_panel<1024>, dim3(batch), dim3(1024), 0, stream, cur, Np, k0, pout, src, piv, perm,
                         pinv);
    TICC_CHECK_LAUNCH();
    size_t slot = 0;
    if (sample) {
      slot = g_prof.used++;
      TICC_CHECK_HIP(hipEventRecord(g_prof.events[2 * slot], stream));
    }
    hipLaunchKernelGGL(k_gj_update, dim3(tiles * tiles, batch), dim3(256), 0, stream, cur, nxt, Np, k0, pout, src,
                       piv);
    TICC_CHECK_LAUNCH();
    if (sample) TICC_CHECK_HIP(hipEventRecord(g_prof.events[2 * slot + 1], stream));
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
