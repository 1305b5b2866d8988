// Batched blocked Gauss-Jordan inversion with partial pivoting (gfx950).
//
// A batch of padded Np x Np FP64 matrices is inverted in Np / NB block steps
// of NB columns, taken two at a time (steps k, k+1 -> one rank-2NB update).
// The panel chain of a pair runs in one workgroup per matrix
// (gj_pair_chain_body):
//   1. panel of step k: in-place Gauss-Jordan on the Np x NB column panel,
//      pivot rows chosen by partial pivoting among the rows not yet used
//      (LAPACK idamax order: largest |a|, first position on ties), the panel
//      in registers, up to 2 rows per thread -> transformed panel P_k in
//      output-row order, row source map s_k, pivot rows p_k;
//   2. Z[i] = mask_k(i) X[s_k(i)][cols_{k+1}] + P_k[i] . X[p_k][cols_{k+1}]
//      (the columns of step k+1 after step k, FP64 MFMA);
//   3. panel of Z -> P_{k+1}, s_{k+1}, p_{k+1}, the combined row map, the 2NB
//      rows of X the pair multiplies and the transposed coefficients;
//   4. A_k[i] = mask_{k+1}(i) P_k[s_{k+1}(i)] + P_{k+1}[i] . P_k[p_{k+1}].
// Then k_gj_update<2NB> applies
//   Out[i][:] = m(i) X[s_k(s_{k+1}(i))][:] + [A_k | P_{k+1}][i] . [X[p_k]; X[s_k(p_{k+1})]]
// -- one pass over X for two Gauss-Jordan steps, FP64 MFMA, reading X and
// writing the ping-pong partner buffer; the panel columns are replaced by the
// coefficients.  An odd last step uses the rank-NB form.  The panel chain of
// pair q+1 runs inside the update of pair q (k_gj_fused); it needs only a
// one-tile look-ahead of that update.
// Row swaps compose into perm (original row at each position); the inverse of
// the matrix is Out with its columns gathered through pinv = perm^-1, which the
// consumers fuse into their own reads.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <utility>
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
  double rv[4];
  int rp[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    rv[r] = readlane_f64(v, 16 * r);
    rp[r] = __builtin_amdgcn_readlane(p, 16 * r);
  }
#pragma unroll
  for (int r = 0; r < 4; r += 2) {
    if (better(rv[r + 1], rp[r + 1], rv[r], rp[r])) {
      rv[r] = rv[r + 1];
      rp[r] = rp[r + 1];
    }
  }
  if (better(rv[2], rp[2], rv[0], rp[0])) {
    rv[0] = rv[2];
    rp[0] = rp[2];
  }
  v = rv[0];
  p = rp[0];
}

// One step of the panel's cross-wave winner reduction: take the record of the
// lane CTRL's row rotation brings when it is better (larger key, smaller
// position on ties).
template <int CTRL>
__device__ __forceinline__ void panel_winner_step(unsigned long long& k, int& p, int& w) {
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(k & 0xffffffffull), CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(k >> 32), CTRL, 0xF, 0xF, false);
  const int op = __builtin_amdgcn_update_dpp(0, p, CTRL, 0xF, 0xF, false);
  const int ow = __builtin_amdgcn_update_dpp(0, w, CTRL, 0xF, 0xF, false);
  const unsigned long long ok = ((unsigned long long)hi << 32) | lo;
  const bool tb = (ok > k) | ((ok == k) & (op < p));
  k = tb ? ok : k;
  p = tb ? op : p;
  w = tb ? ow : w;
}

template <typename T>
__device__ __forceinline__ T readlane_t(T v, int lane);
template <>
__device__ __forceinline__ double readlane_t<double>(double v, int lane) {
  return readlane_f64(v, lane);
}
template <>
__device__ __forceinline__ float readlane_t<float>(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// A wave's pivot candidate in the panel's column step (32 bytes).
// The one-barrier column schedule also carries the candidate's multiplier of
// the previous pivot (gprev).
template <typename T, bool GPREV>
struct alignas(16) PanelRec {
  unsigned long long key;
  int pos, row;
  T pinv, pnext, gprev;
};
template <typename T>
struct alignas(16) PanelRec<T, false> {
  unsigned long long key;
  int pos, row;
  T pinv, pnext;
};

// Panel: one workgroup per matrix; thread t owns rows t + THREADS q
// (q < RPT, Np <= THREADS RPT) with their NB panel values in registers and
// its rows' current positions.  Per column: each wave reduces its best
// candidate (largest |a|, first position on ties) and its winning lane
// publishes a record of it in a double-buffered LDS slot; the column's
// barrier publishes these records and the previous column's pivot row.  After
// it every thread performs the previous pivot's in-place Gauss-Jordan update of
// the other columns, streaming that pivot row from LDS, picks the global
// winner, applies the row swap to its own position records and updates the
// next column, while the winner's owner writes its row to LDS.  (With three
// rows per thread a second barrier per column publishes the pivot row and the
// other columns take the current pivot at once.)  Optionally the panel also leaves its row map and pivot rows
// in LDS (lsrc, lpiv) and reads the previous panel's there (lprev_src,
// lprev_piv) instead of the global copies.
template <int THREADS, int RPT, typename T>
__device__ __forceinline__ void gj_panel_body(const T* __restrict__ X, long long bstride, int ld, int col0,
                                              int Np, int k0, T* __restrict__ Pout, int* __restrict__ src,
                                              int* __restrict__ piv, int* __restrict__ perm, int* __restrict__ pinv,
                                              T* __restrict__ PoutT, long long tstride,
                                              const int* __restrict__ prev_src, const int* __restrict__ prev_piv,
                                              int prev_k0, int* __restrict__ src2, int* __restrict__ brow, int b,
                                              int* lsrc = nullptr, int* lpiv = nullptr,
                                              const int* lprev_src = nullptr, const int* lprev_piv = nullptr) {
  typedef Elem<T> E;
  constexpr int NW = THREADS / 64;
  static_assert((THREADS & (THREADS - 1)) == 0, "the panel's thread count is a power of two");
  constexpr bool ONE_BARRIER = RPT <= 2;  // (the column schedule, below)
  __shared__ PanelRec<T, ONE_BARRIER> s_rec[2][NW];
  __shared__ T s_prow[2][NB];
  __shared__ T s_inv[NB];

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const T* Xb = X + (size_t)b * bstride;

  T v[RPT][NB];
  int pos[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = t + THREADS * q;
    if (r < Np) {
      const typename E::vec2* rp = reinterpret_cast<const typename E::vec2*>(Xb + (size_t)r * ld + col0);
#pragma unroll
      for (int l = 0; l < NB / 2; ++l) {
        const typename E::vec2 d = rp[l];
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

  // The column loop runs PU columns per iteration on a rotated register
  // window: v[q][l] holds panel column (j0 + l) % NB, so every iteration's
  // code is the same and the kernel stays small enough for the instruction
  // cache (a fully unrolled 32-column loop does not fit and runs from L2).
  //
  // Up to two rows per thread, one barrier per column: the update of the
  // other columns by pivot j - 1 runs after column j's barrier (which
  // publishes pivot row j - 1 with the candidate records).  A candidate's value
  // in column j + 1 is then current through pivot j - 2 only, so its record
  // also carries the row's pivot-(j - 1) multiplier, and the winner's value is
  // brought through pivot j - 1 by the same FMA the row's own update applies:
  // every value is formed by the same operations as with a second barrier per
  // column.  (With three rows per thread the registers the deferred update
  // keeps live make the panel spill: 8 x 1100 inversions took 9.0 against
  // 7.0 ms.  There a second barrier per column publishes the pivot row and the
  // other columns take it at once.)
  constexpr int PU = 4;
  T g[RPT];     // the multipliers of the previous column's pivot
  int ppr = -1;  // the previous column's pivot row
#pragma unroll
  for (int q = 0; q < RPT; ++q) g[q] = T(0);
  // Pivot (j - 1)'s update of the other columns at column step u of the
  // window: its column is u - 1, or NB - 1 after the window's rotation at
  // u = 0, where its pivot row in LDS still has the previous window's
  // coordinates (index + PU); column u took it already, as the next column.
  auto prev_update = [&](int u, const T* prow) {
    const int lp = (u == 0) ? NB - 1 : u - 1;
    const int sh = (u == 0) ? PU : 0;
#pragma unroll
    for (int l = 0; l < NB; ++l) {
      if (l == u || l == lp) continue;
      const T pl = prow[(l + sh) % NB];
#pragma unroll
      for (int q = 0; q < RPT; ++q) v[q][l] = E::fma(-g[q], pl, v[q][l]);
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) v[q][lp] = (t + THREADS * q == ppr) ? T(1) : -g[q];
  };
  // The same update by the current column's pivot (two-barrier schedule).
  auto cur_update = [&](int u, const T* prow, int pr) {
#pragma unroll
    for (int l = 0; l < NB; ++l) {
      if (l == u || l == u + 1) continue;
      const T pl = prow[l];
#pragma unroll
      for (int q = 0; q < RPT; ++q) v[q][l] = E::fma(-g[q], pl, v[q][l]);
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) v[q][u] = (t + THREADS * q == pr) ? T(1) : -g[q];
  };
  for (int j0 = 0; j0 < NB; j0 += PU) {
#pragma unroll
  for (int u = 0; u < PU; ++u) {
    const int j = j0 + u;
    const int c = k0 + j;
    const int buf = u & 1;
    // This thread's candidate: largest |a| among its rows at positions >= c,
    // first position on ties (NaN ranks as +inf so a column always has a pivot
    // and the NaN propagates to the result).
    // (bitwise & and |: no short-circuit branches around the comparisons)
    double cv = -1.0;
    int cp = 0x7fffffff;
    int cq = -1;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      double a = fabs((double)v[q][u]);
      a = (a != a) ? __builtin_inf() : a;
      const bool take = (pos[q] >= c) & ((a > cv) | ((a == cv) & (pos[q] < cp)));
      cv = take ? a : cv;
      cp = take ? pos[q] : cp;
      cq = take ? q : cq;
    }
    // The candidate's signed value, its value in the next column and the
    // reciprocal of the value, formed by every lane before the wave's winner is
    // known, so that the division overlaps the reductions (only the winning
    // lane's are published).
    T pv = v[0][u], pnx = v[0][u + 1], gpv = ONE_BARRIER ? g[0] : T(0);
#pragma unroll
    for (int q = 1; q < RPT; ++q) {
      pv = (cq == q) ? v[q][u] : pv;
      pnx = (cq == q) ? v[q][u + 1] : pnx;
      if constexpr (ONE_BARRIER) gpv = (cq == q) ? g[q] : gpv;
    }
    const T rinv = T(1) / pv;
    // Exact ordering key of the candidate: the bits of |a| plus one (monotone in
    // |a|), 0 for no candidate.  The wave maximum of its top 32 bits takes one
    // DPP max per step; only when several lanes share that maximum is the exact
    // (|a|, position) argmax run among the wave.
    const unsigned long long key = (cq >= 0) ? (unsigned long long)__double_as_longlong(cv) + 1ull : 0ull;
    const unsigned khi = (unsigned)(key >> 32);
    unsigned m = khi;
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x128, 0xF, 0xF, false));  // row_ror:8
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x124, 0xF, 0xF, false));  // row_ror:4
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x122, 0xF, 0xF, false));  // row_ror:2
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x121, 0xF, 0xF, false));  // row_ror:1
    // the four rows' maxima into lane 63: row_bcast:15 into rows 1 and 3, then
    // row_bcast:31 into rows 2 and 3 (lanes of the other rows take 0)
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x142, 0xA, 0xF, false));
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x143, 0xC, 0xF, false));
    const unsigned M = (unsigned)__builtin_amdgcn_readlane((int)m, 63);
    // A lane has a candidate whenever cq >= 0, even when |a| is zero or
    // subnormal (its key then lies below 2^32 and khi is 0): a column of exact
    // zeros below the diagonal still yields a pivot (value 0, so the inverse is
    // non-finite, as LAPACK reports a singular matrix) and never an empty
    // selection.
    int wl = -1;
    if (__ballot(cq >= 0) != 0ull) {
      const bool top = cq >= 0 && khi == M;
      const unsigned long long tied = __ballot(top);
      if (__popcll(tied) == 1) {
        wl = __ffsll((long long)tied) - 1;
      } else {
        double wv = top ? cv : -1.0;
        int wp = top ? cp : 0x7fffffff;
        wave_argmax_dpp(wv, wp);
        wl = __ffsll((long long)__ballot(top && cv == wv && cp == wp)) - 1;
      }
    }
    // The wave's winning lane publishes its candidate's record: key, position,
    // row, the reciprocal of its pivot value, its value in the next column and
    // its multiplier of the previous pivot (the row itself is written once the
    // global winner is known).
    if (lane == wl) {
      PanelRec<T, ONE_BARRIER> rec;
      rec.key = key;
      rec.pos = cp;
      rec.row = t + THREADS * cq;
      rec.pinv = rinv;
      rec.pnext = pnx;
      if constexpr (ONE_BARRIER) rec.gprev = gpv;
      s_rec[buf][wave] = rec;
    } else if (wl < 0 && lane == 0) {
      PanelRec<T, ONE_BARRIER> rec;
      rec.key = 0ull;
      rec.pos = 0x7fffffff;
      rec.row = 0;
      rec.pinv = T(0);
      rec.pnext = T(0);
      if constexpr (ONE_BARRIER) rec.gprev = T(0);
      s_rec[buf][wave] = rec;
    }
    __syncthreads();
    if (ONE_BARRIER && j > 0) prev_update(u, s_prow[buf ^ 1]);
    // Global winner: larger key, smaller position on ties.  Lane l reads wave
    // (l mod NW)'s record, so every NW consecutive lanes of a 16-lane row hold
    // all records, and log2(NW) DPP row rotations leave the winner in every
    // lane; its row, reciprocal and next-column value are then read from lane
    // bw, which loaded wave bw's record.  (A wave without a candidate has key 0
    // and never wins: some wave always has a candidate.)
    const int wr = lane & (NW - 1);
    const PanelRec<T, ONE_BARRIER> rr = s_rec[buf][wr];
    const int rrow = rr.row;
    const T rpinv = rr.pinv;
    const T rpn = rr.pnext;
    T rgp = T(0);
    if constexpr (ONE_BARRIER) rgp = rr.gprev;
    // First the top 32 bits of the keys with their lowest log2(NW) bits
    // replaced by the wave (smaller wave ranks higher): when only one wave's key
    // reaches the maximum there, that wave's full key is larger than every
    // other and it is the winner; otherwise the exact (key, position) reduction
    // decides.
    const unsigned wkhi = (unsigned)(rr.key >> 32) & ~(unsigned)(NW - 1);
    unsigned cm = wkhi | (unsigned)(NW - 1 - wr);
    cm = max(cm, (unsigned)__builtin_amdgcn_update_dpp(0, (int)cm, 0x121, 0xF, 0xF, false));  // row_ror:1
    if constexpr (NW > 2) cm = max(cm, (unsigned)__builtin_amdgcn_update_dpp(0, (int)cm, 0x122, 0xF, 0xF, false));
    if constexpr (NW > 4) cm = max(cm, (unsigned)__builtin_amdgcn_update_dpp(0, (int)cm, 0x124, 0xF, 0xF, false));
    const unsigned cmax = (unsigned)__builtin_amdgcn_readfirstlane((int)cm);
    const unsigned long long level = __ballot(lane < NW && wkhi == (cmax & ~(unsigned)(NW - 1)));
    int bw, bp;
    if (__popcll(level) == 1) {
      bw = NW - 1 - (int)(cmax & (unsigned)(NW - 1));
      bp = __builtin_amdgcn_readlane(rr.pos, bw);
    } else {
      unsigned long long kb = rr.key;
      int pb = rr.pos;
      int wb = wr;
      panel_winner_step<0x121>(kb, pb, wb);  // row_ror:1
      if constexpr (NW > 2) panel_winner_step<0x122>(kb, pb, wb);  // row_ror:2
      if constexpr (NW > 4) panel_winner_step<0x124>(kb, pb, wb);  // row_ror:4
      bw = __builtin_amdgcn_readfirstlane(wb);
      bp = __builtin_amdgcn_readfirstlane(pb);
    }
    const int pr = __builtin_amdgcn_readlane(rrow, bw);
    const T inv = readlane_t<T>(rpinv, bw);
    T pn = readlane_t<T>(rpn, bw);
    // (the winner's next-column value through pivot j - 1, as its own row
    // update formed it)
    if (ONE_BARRIER && j > 0)
      pn = E::fma(-readlane_t<T>(rgp, bw), s_prow[buf ^ 1][(u + 1 + (u == 0 ? PU : 0)) % NB], pn);
    // The pivot row's owner writes it for every thread; the next column's
    // barrier (or the second barrier of this one) publishes it.
    if (t == (pr & (THREADS - 1))) {
      const int pq = pr / THREADS;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        if (q == pq) {
#pragma unroll
          for (int l = 0; l < NB; ++l) s_prow[buf][l] = v[q][l];
        }
      }
    }
    // Gauss-Jordan step with the pivot row's scaling deferred: every other row
    // r takes a_rl -= (a_rj / p) p_l, a_rj = -a_rj / p; the pivot row keeps its
    // values with a_rj = 1.  A row that was a pivot row stays p times its true
    // values through the later columns of the panel (the update is linear in
    // the row) and is multiplied by 1/p (kept in s_inv) once at the end.  The
    // pivot row's own update is a no-op through g = 0, so no lane branches.
    // The next column is updated first: the next pivot search needs only it.
    if (t == 0) s_inv[j] = inv;
    ppr = pr;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const bool is_piv = (t + THREADS * q == pr);
      if (pos[q] == c) pos[q] = bp;
      pos[q] = is_piv ? c : pos[q];
      g[q] = is_piv ? T(0) : v[q][u] * inv;
    }
    // (u + 1 is the next column, or for the last column another column of
    // the window; pn is the pivot row's value there either way)
#pragma unroll
    for (int q = 0; q < RPT; ++q) v[q][u + 1] = E::fma(-g[q], pn, v[q][u + 1]);
    if constexpr (!ONE_BARRIER) {
      __syncthreads();
      cur_update(u, s_prow[buf], pr);
    }
    }
    // rotate the window by PU columns
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      T tmp[PU];
#pragma unroll
      for (int l = 0; l < PU; ++l) tmp[l] = v[q][l];
#pragma unroll
      for (int l = 0; l + PU < NB; ++l) v[q][l] = v[q][l + PU];
#pragma unroll
      for (int l = 0; l < PU; ++l) v[q][NB - PU + l] = tmp[l];
    }
  }

  // The last pivot's update of the other columns (one-barrier schedule; the
  // window is back at column 0), then the deferred pivot-row scaling of this
  // panel's pivot rows.
  __syncthreads();
  if constexpr (ONE_BARRIER) prev_update(0, s_prow[(NB - 1) & 1]);
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    if (pos[q] >= k0 && pos[q] < k0 + NB) {
      const T ri = s_inv[pos[q] - k0];
#pragma unroll
      for (int l = 0; l < NB; ++l) v[q][l] *= ri;
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
    typename E::vec2* op = reinterpret_cast<typename E::vec2*>(Pout + ((size_t)b * Np + p) * NB);
#pragma unroll
    for (int l = 0; l < NB / 2; ++l) op[l] = E::pair(v[q][2 * l], v[q][2 * l + 1]);
    src[(size_t)b * Np + p] = r;
    const bool blk = (p >= k0 && p < k0 + NB);
    if (blk) piv[b * NB + (p - k0)] = r;
    if (lsrc) lsrc[p] = r;
    if (blk && lpiv) lpiv[p - k0] = r;
    if (PoutT) {
#pragma unroll
      for (int l = 0; l < NB; ++l) PoutT[(size_t)b * tstride + (size_t)l * Np + p] = v[q][l];
    }
    if (src2) {  // row maps of the update that follows this panel
      if (prev_src) {  // second step of a pair: compose with the first step's maps
        const int s1r = lprev_src ? lprev_src[r] : prev_src[(size_t)b * Np + r];
        src2[(size_t)b * Np + p] = (blk || (r >= prev_k0 && r < prev_k0 + NB)) ? -1 : s1r;
        if (blk) brow[b * 2 * NB + NB + (p - k0)] = s1r;
      } else {
        src2[(size_t)b * Np + p] = blk ? -1 : r;
        if (blk) brow[b * 2 * NB + (p - k0)] = r;
      }
    }
    pold[q] = (k0 == 0) ? r : perm[(size_t)b * Np + r];
  }
  if (src2 && prev_src && t < NB) brow[b * 2 * NB + t] = lprev_piv ? lprev_piv[t] : prev_piv[b * NB + t];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    if (pos[q] < 0) continue;
    perm[(size_t)b * Np + pos[q]] = pold[q];
    if (k0 + NB >= Np) pinv[(size_t)b * Np + pold[q]] = pos[q];
  }
}

template <int THREADS, int RPT, typename T>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(THREADS / 256, THREADS / 256))) void k_gj_panel(
    const T* __restrict__ X, long long bstride, int ld, int col0, int Np, int k0, T* __restrict__ Pout,
    int* __restrict__ src, int* __restrict__ piv, int* __restrict__ perm, int* __restrict__ pinv,
    T* __restrict__ PoutT, long long tstride, const int* __restrict__ prev_src, const int* __restrict__ prev_piv,
    int prev_k0, int* __restrict__ src2, int* __restrict__ brow, const int* __restrict__ mask) {
  if (mask && !mask[blockIdx.x]) return;
  gj_panel_body<THREADS, RPT, T>(X, bstride, ld, col0, Np, k0, Pout, src, piv, perm, pinv, PoutT, tstride, prev_src,
                              prev_piv, prev_k0, src2, brow, blockIdx.x);
}

// Raw buffer loads: an offset past the end of the buffer reads 0 without a
// memory access, so masked rows need no branch (a branch around a load makes
// the compiler wait for that load at the join).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
constexpr int kOutOfRange = 0x7ff00000;
__device__ __forceinline__ double buffer_load_f64(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ double2 buffer_load_f64x2(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
// One element / a pair of elements of type T at byte offset `off`.
template <typename T>
__device__ __forceinline__ T load1(__amdgpu_buffer_rsrc_t r, int off);
template <>
__device__ __forceinline__ double load1<double>(__amdgpu_buffer_rsrc_t r, int off) {
  return buffer_load_f64(r, off);
}
template <>
__device__ __forceinline__ float load1<float>(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <typename T>
__device__ __forceinline__ typename Elem<T>::vec2 load2(__amdgpu_buffer_rsrc_t r, int off);
template <>
__device__ __forceinline__ double2 load2<double>(__amdgpu_buffer_rsrc_t r, int off) {
  return buffer_load_f64x2(r, off);
}
template <>
__device__ __forceinline__ float2 load2<float>(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// Cache policy of the update's streamed X reads and Out writes: both are
// marked non-temporal (buffer-load cache-policy bit 1, nontemporal stores), so
// the matrices streamed through a launch, each read and written once, displace
// less of what is re-read from L2 -- the staged B rows and the panel chains'
// panels, row maps and coefficients.  At N = 800 x 64 energies the fused launch
// went from 232.8 to 223.2 us and the sweep from 8355 to 8698 energy.R-steps/s
// (loads alone 8505, stores alone 8610); results are bit for bit unchanged.
// Compile-time switches: 0 restores the default policy.
#ifndef TICC_UPD_LOAD_CPOL
#define TICC_UPD_LOAD_CPOL 2
#endif
#ifndef TICC_UPD_STORE_NT
#define TICC_UPD_STORE_NT 1
#endif
template <typename T, int CPOL>
__device__ __forceinline__ T load1_cpol(__amdgpu_buffer_rsrc_t r, int off) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, CPOL));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, CPOL));
}

// Small panel GEMM (FP64 MFMA) of the paired block step, for one matrix b:
//   Out[i][0:NB] = mask(i) C[gC(i)][colC : colC+NB] + A[i][0:NB] . B[gB(r)][colB : colB+NB],
//   mask(i) = 0 for i in [mlo, mlo+NB), 1 otherwise.
// Out is row-major (Np x NB) or, with `ostride` > 0, transposed (NB rows of
// length Np, matrix stride ostride).  The row maps gC (Np entries) and gB
// (NB entries) are the panel's, in this workgroup's LDS.  Wave w takes the
// 16-row blocks w, w + THREADS/64, ...; the C and A operands of GCH blocks are
// loaded together, so the gathers' latency is paid once per GCH blocks instead
// of once per block.  sB stages the NB gathered B rows.
template <int THREADS, int MAXROWS, typename T>
__device__ __forceinline__ void panel_gemm_body(const T* __restrict__ A, const T* __restrict__ C,
                                                long long cb, int ldC, int colC, const int* gC,
                                                const T* __restrict__ Bm, long long bb, int ldB, int colB,
                                                const int* gB, int mlo, int Np,
                                                T* __restrict__ Out, long long ostride, int b,
                                                T (*sB)[NB + 1]) {
  typedef Elem<T> E;
  constexpr int RB = THREADS / 4;                 // rows per round of all waves
  constexpr int MAXIT = (MAXROWS + RB - 1) / RB;  // rounds at the largest Np of the caller
  constexpr int GCH = 4;                          // rounds whose operands are in flight together
  static_assert(MAXIT % GCH == 0, "round count must be a multiple of the chunk");
  const int t = threadIdx.x;
  const int wave = t >> 6, lane = t & 63, lr = lane & 15, lq = lane >> 4;
  const T* Cb = C + (size_t)b * cb;
  const auto rC = make_rsrc(Cb, (unsigned)(sizeof(T) * Np * ldC));
  const auto rA = make_rsrc(A + (size_t)b * Np * NB, (unsigned)(sizeof(T) * Np * NB));
  const T* Bb = Bm + (size_t)b * bb;
  for (int e = t; e < NB * NB; e += THREADS) {
    const int r = e / NB, cc = e % NB;
    sB[r][cc] = Bb[(size_t)gB[r] * ldB + colB + cc];
  }
  __syncthreads();
  for (int c0 = 0; c0 < MAXIT; c0 += GCH) {
    if (c0 * RB >= Np) break;
    // source row of each accumulator row (-1: masked or past the end)
    int src[GCH][4];
#pragma unroll
    for (int u = 0; u < GCH; ++u)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = (c0 + u) * RB + wave * 16 + E::drow(lq, r);
        const bool live = row < Np && !(row >= mlo && row < mlo + NB);
        src[u][r] = live ? gC[row] : -1;
      }
    // (the products are accumulated from zero and C added once at the end, as
    // in the update: an accumulator started at C would take the NB products one
    // at a time onto it)
    typename E::acc cvals[GCH][2];
    T a[GCH][NB / 4];
#pragma unroll
    for (int u = 0; u < GCH; ++u) {
      const int row0 = (c0 + u) * RB + wave * 16;
#pragma unroll
      for (int kk = 0; kk < NB / 4; ++kk)
        a[u][kk] = load1<T>(rA, row0 + lr < Np ? (int)sizeof(T) * ((row0 + lr) * NB + 4 * kk + lq) : kOutOfRange);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int s = src[u][r];
        const int base = s >= 0 ? (int)sizeof(T) * (s * ldC + colC + lr) : kOutOfRange;
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) cvals[u][bj][r] = load1<T>(rC, base + 16 * (int)sizeof(T) * bj);
      }
    }
#pragma unroll
    for (int u = 0; u < GCH; ++u) {
      const int row0 = (c0 + u) * RB + wave * 16;
      typename E::acc acc[2];
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[bj][r] = T(0);
#pragma unroll
      for (int kk = 0; kk < NB / 4; ++kk) {
        const int kq = 4 * kk + lq;
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
          acc[bj] = E::mfma(a[u][kk], sB[kq][16 * bj + lr], acc[bj]);
      }
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[bj][r] = cvals[u][bj][r] + acc[bj][r];
      if (ostride > 0) {
        T* Ob = Out + (size_t)b * ostride;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = row0 + E::drow(lq, r);
          if (row < Np) {
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) Ob[(size_t)(16 * bj + lr) * Np + row] = acc[bj][r];
          }
        }
      } else {
        T* Ob = Out + (size_t)b * Np * NB;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = row0 + E::drow(lq, r);
          if (row < Np) {
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) Ob[(size_t)row * NB + 16 * bj + lr] = acc[bj][r];
          }
        }
      }
    }
  }
}

// The whole panel chain of one pair of steps (k0, k0+NB) for matrix b, by one
// workgroup: panel of step k0 read from the columns [xcol, xcol+NB) of Xc;
// Z = the columns of step k0+NB after step k0; panel of Z (also writing the
// transposed coefficients and the pair's row maps); the combined step-k0
// coefficients, written transposed to AT[0:NB].
template <typename T>
struct PairJob {
  const T* Xc;
  long long xs;
  int ldx, xcol, Np, k0;
  int single;  // 1: only the panel of the trailing single step (with its update maps)
  T *P1, *P2, *Z, *AT;
  int *s1, *s2, *p1, *p2, *src2, *brow, *perm, *pinv;
  const int* mask;  // null, or 0 for the matrices this inversion skips
};

template <int RPT, typename T>
__device__ __forceinline__ void gj_pair_chain_body(const PairJob<T>& P, int b) {
  // row bound of the panel GEMMs: 1024 up to two rows per thread, else MAX_NP
  constexpr int GR = RPT <= 2 ? 1024 : MAX_NP;
  __shared__ T sG[NB][NB + 1];
  // each panel's row map and pivot rows, also kept here for the GEMM after it
  // and the second panel (the global copies serve the update)
  __shared__ int s_map[2][GR];
  __shared__ int s_pv[2][NB];
  const int Np = P.Np, k0 = P.k0, k1 = P.k0 + NB;
  const long long pn = (long long)Np * NB;
  const long long at_stride = 2LL * NB * Np;
  if (P.single) {
    gj_panel_body<512, RPT, T>(P.Xc, P.xs, P.ldx, P.xcol, Np, k0, P.P1, P.s1, P.p1, P.perm, P.pinv, P.AT, at_stride,
                            nullptr, nullptr, 0, P.src2, P.brow, b);
    return;
  }
  gj_panel_body<512, RPT, T>(P.Xc, P.xs, P.ldx, P.xcol, Np, k0, P.P1, P.s1, P.p1, P.perm, P.pinv, nullptr, 0, nullptr,
                          nullptr, 0, nullptr, nullptr, b, s_map[0], s_pv[0]);
  __syncthreads();
  panel_gemm_body<512, GR, T>(P.P1, P.Xc, P.xs, P.ldx, P.xcol + NB, s_map[0], P.Xc, P.xs, P.ldx, P.xcol + NB, s_pv[0], k0,
                              Np, P.Z, 0LL, b, sG);
  __syncthreads();
  gj_panel_body<512, RPT, T>(P.Z, pn, NB, 0, Np, k1, P.P2, P.s2, P.p2, P.perm, P.pinv, P.AT + (size_t)NB * Np, at_stride,
                          P.s1, P.p1, k0, P.src2, P.brow, b, s_map[1], s_pv[1], s_map[0], s_pv[0]);
  __syncthreads();
  panel_gemm_body<512, GR, T>(P.P2, P.P1, pn, NB, 0, s_map[1], P.P1, pn, NB, 0, s_pv[1], k1, Np, P.AT, at_stride, b, sG);
}

template <int RPT, typename T>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_gj_pair_chain(PairJob<T> P) {
  if (P.mask && !P.mask[blockIdx.x]) return;
  gj_pair_chain_body<RPT, T>(P, blockIdx.x);
}

// Rank-KB update (KB = NB for one step, 2 NB for a pair of steps):
//   Out[i][:] = (src2(i) < 0 ? 0 : X[src2(i)][:]) + AT[:, i]^T . X[brow][:],
// the columns [k0, k0+KB) replaced by AT^T.  Only the 64-column tiles [T0, T1)
// are produced; column j lands in Out[i * ldo + j - ocol0] (a full update has
// ldo = Np, ocol0 = 0; the look-ahead of the next pair's panel columns writes
// one tile into an Np x 2NB buffer).  AT holds the KB coefficient columns
// transposed (KB rows of length Np, matrix stride 2 NB Np) so an MFMA A fragment
// is 16 consecutive doubles.  A workgroup of NBANDS x 4 waves (32 x 32 each)
// walks NBANDS adjacent 64-row bands across a group of 64-column tiles, sharing
// each staged B tile between the bands: the bands' A fragments stay in
// registers; a tile's C rows are loaded before its MFMAs (which accumulate the
// products from zero; C is added once at the end) and the next tile's B rows
// (into the idle LDS buffer) while the current tile's MFMAs and stores run.
template <int KB, int NBANDS, bool ONE, typename T>
__device__ __forceinline__ void gj_update_body(const T* __restrict__ X, T* __restrict__ Out,
                                               long long ostride, int ldo, int ocol0, int Np, int k0,
                                               const T* __restrict__ AT, const int* __restrict__ src2,
                                               const int* __restrict__ brow, int b, int band0, int jt0, int jt1,
                                               int t, T (*sB)[KB][64]) {
  typedef Elem<T> E;
  constexpr int SZ = (int)sizeof(T);
  constexpr int KK = KB / 4;
  constexpr int RP = 8 * NBANDS;  // B-tile rows staged per pass by the whole workgroup
  const int th = t & 255;         // thread within its band's 4 waves
  const int band = band0 + (NBANDS > 1 ? (t >> 8) : 0);
  const int wave = th >> 6, lane = th & 63, lr = lane & 15, lq = lane >> 4;
  const int i0 = band * 64 + (wave >> 1) * 32;
  const int jw = (wave & 1) * 32;
  const bool rows_ok = i0 < Np;
  const T* Xb = X + (size_t)b * Np * Np;
  T* Ob = Out + (size_t)b * ostride - ocol0;
  const auto rX = make_rsrc(Xb, (unsigned)(sizeof(T) * Np * Np));
  const auto rA = make_rsrc(AT + (size_t)b * 2 * NB * Np, (unsigned)(sizeof(T) * KB * Np));
  const auto rS = make_rsrc(src2 + (size_t)b * Np, (unsigned)(sizeof(int) * Np));
  const int* brb = brow + b * 2 * NB;

  T a[2][KK];
  int sr[2][4];
#pragma unroll
  for (int kk = 0; kk < KK; ++kk)
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
      a[bi][kk] = load1<T>(rA, rows_ok ? SZ * ((4 * kk + lq) * Np + i0 + 16 * bi + lr) : kOutOfRange);
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int v = (int)__builtin_amdgcn_raw_buffer_load_b32(
          rS, rows_ok ? 4 * (i0 + 16 * bi + E::drow(lq, r)) : kOutOfRange, 0, 0);
      sr[bi][r] = rows_ok ? v : -1;
    }
  const int bk = t >> 5, c2 = (t & 31) * 2;
  int bro[KB / RP];
#pragma unroll
  for (int h = 0; h < KB / RP; ++h) bro[h] = brb[h * RP + bk];

  // 0: update tile, 1: coefficient (panel) columns, 2: outside the matrix
  auto kind = [&](int J) {
    const int j0 = J * 64 + jw;
    if (!rows_ok || j0 >= Np) return 2;
    return (j0 >= k0 && j0 < k0 + KB) ? 1 : 0;
  };
  // live = false: every load takes the out-of-range offset (reads 0, no
  // memory access).  The B prefetch is issued on every path so the compiler's
  // wait counts before the current tile's MFMAs need not cover it.
  auto load_c = [&](int J, bool live, typename E::acc(&c)[2][2]) {
    const int j0 = J * 64 + jw;
    const bool upd = live && kind(J) == 0;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int s = sr[bi][r];
        const int base = (upd && s >= 0) ? SZ * (s * Np + j0 + lr) : kOutOfRange;
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) c[bi][bj][r] = load1_cpol<T, TICC_UPD_LOAD_CPOL>(rX, base + 16 * SZ * bj);
      }
  };
  typename E::vec2 bst[KB / RP];
  auto load_b = [&](int J, bool live) {
    const bool in = live && (J * 64 + c2) < Np;
#pragma unroll
    for (int h = 0; h < KB / RP; ++h)
      bst[h] = load2<T>(rX, in ? SZ * (bro[h] * Np + J * 64 + c2) : kOutOfRange);
  };
  // B tile rows are stored with bit 4 of the column flipped on odd rows, so the
  // four rows an MFMA B fragment reads fall in different LDS bank halves.
  auto store_b = [&](int buf) {
#pragma unroll
    for (int h = 0; h < KB / RP; ++h) {
      const int k = h * RP + bk;
      *reinterpret_cast<typename E::vec2*>(&sB[buf][k][c2 ^ ((k & 1) << 4)]) = bst[h];
    }
  };
  // The products are accumulated from zero and the tile's C values added once
  // at the end: an accumulator that started at C would take the KB products
  // one at a time onto the much larger C (the MFMA adds its four products
  // sequentially), losing in a biased way the parts of small same-signed
  // products below C's last bit (the inverse of I - h^2 T/6 came out with a
  // mean diagonal error of -6 ulp).  C's loads are issued before the tile's
  // MFMAs, which hide their latency.
  auto compute_store = [&](int J, int buf, const typename E::acc(&cv)[2][2]) {
    const int kd = kind(J);
    const int j0 = J * 64 + jw;
    if (kd == 0) {
      typename E::acc acc[2][2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[bi][bj][r] = T(0);
#pragma unroll
      for (int kk = 0; kk < KK; ++kk) {
        const int kq = 4 * kk + lq;
        const int sw = (kq & 1) << 4;
        T bv[2];
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) bv[bj] = sB[buf][kq][(jw + 16 * bj + lr) ^ sw];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
          for (int bj = 0; bj < 2; ++bj)
            acc[bi][bj] = E::mfma(a[bi][kk], bv[bj], acc[bi][bj]);
      }
      // Stores stream past the caches, except the one-tile look-ahead's, which
      // the next panel chain reads right away.  (Keeping a pair update's tile
      // that the next look-ahead reads cached as well was slower.)
      const bool keep = ONE;
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          T* orow = Ob + (size_t)(i0 + 16 * bi + E::drow(lq, r)) * ldo + j0 + lr;
#pragma unroll
          for (int bj = 0; bj < 2; ++bj) {
            const T v = cv[bi][bj][r] + acc[bi][bj][r];
            if (TICC_UPD_STORE_NT && !keep) __builtin_nontemporal_store(v, orow + 16 * bj);
            else orow[16 * bj] = v;
          }
        }
    } else if (kd == 1) {  // coefficient columns: Out[i][j0 + c] = AT[j0 - k0 + c][i]
      const int kb = j0 - k0;
      T pv[2][4][2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int bj = 0; bj < 2; ++bj)
            pv[bi][r][bj] = load1<T>(rA, SZ * ((kb + 16 * bj + lr) * Np + i0 + 16 * bi + E::drow(lq, r)));
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          T* orow = Ob + (size_t)(i0 + 16 * bi + E::drow(lq, r)) * ldo + j0 + lr;
#pragma unroll
          for (int bj = 0; bj < 2; ++bj) orow[16 * bj] = pv[bi][r][bj];
        }
    }
  };

  typename E::acc cv[2][2];
  load_b(jt0, true);
  store_b(0);
  __syncthreads();
  if constexpr (ONE) {  // a single column tile: nothing to prefetch, one B buffer
    load_c(jt0, true, cv);
    compute_store(jt0, 0, cv);
  } else {
  for (int J = jt0;; J += 2) {
    // even tile: LDS buffer 0; the next tile's B rows are loaded into registers
    const bool more1 = J + 1 < jt1;
    load_c(J, true, cv);
    load_b(J + 1, more1);
    compute_store(J, 0, cv);
    if (!more1) break;
    store_b(1);
    __syncthreads();
    // odd tile: LDS buffer 1
    const bool more2 = J + 2 < jt1;
    load_c(J + 1, true, cv);
    load_b(J + 2, more2);
    compute_store(J + 1, 1, cv);
    if (!more2) break;
    store_b(0);
    __syncthreads();
  }
  }
}

// Column tiles [jt0, jt1) of group `grp` out of `ngroups` over [T0, T1).
__device__ __forceinline__ void group_tiles(int T0, int T1, int grp, int ngroups, int& jt0, int& jt1) {
  jt0 = T0 + (grp * (T1 - T0)) / ngroups;
  jt1 = T0 + ((grp + 1) * (T1 - T0)) / ngroups;
}

// (matrix b, workgroup w within it) of a launch with grid (per matrix, matrices).
// Workgroups are dealt to the 8 XCDs round robin in linear order; with a
// matrix count that is a multiple of 8 all workgroups of one matrix share an
// XCD, whose L2 then serves the B rows they all stage.
__device__ __forceinline__ void xcd_matrix_map(int& b, int& w) {
  const int per = gridDim.x;
  const int lin = blockIdx.x + per * blockIdx.y;
  if ((gridDim.y & 7) == 0) {
    const int k = lin >> 3;
    b = (lin & 7) + 8 * (k / per);
    w = k % per;
  } else {
    b = blockIdx.y;
    w = blockIdx.x;
  }
}

// NBANDS = 1: 256 threads, one band per workgroup; NBANDS = 2: 512 threads, two
// bands sharing each B tile (half the B-row traffic).
template <int KB, int NBANDS, typename T>
__global__ __launch_bounds__(256 * NBANDS) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_gj_update(
    const T* __restrict__ X, T* __restrict__ Out, long long ostride, int ldo, int ocol0, int Np, int k0,
    const T* __restrict__ AT, const int* __restrict__ src2, const int* __restrict__ brow, int T0, int T1,
    int ngroups, const int* __restrict__ mask) {
  __shared__ T sB[2][KB][64];
  int b, w;
  xcd_matrix_map(b, w);
  if (mask && !mask[b]) return;
  int jt0, jt1;
  group_tiles(T0, T1, w % ngroups, ngroups, jt0, jt1);
  if (jt0 >= jt1) return;
  gj_update_body<KB, NBANDS, false, T>(X, Out, ostride, ldo, ocol0, Np, k0, AT, src2, brow, b, NBANDS * (w / ngroups), jt0,
                             jt1, threadIdx.x, sB);
}

// One 64-column tile per workgroup (the look-ahead of the next pair's columns):
// without the next tile's prefetch registers and B buffer the kernel fits four
// waves per SIMD, so more of its latency-bound workgroups run at once.
template <int KB, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_gj_update_one(
    const T* __restrict__ X, T* __restrict__ Out, long long ostride, int ldo, int ocol0, int Np, int k0,
    const T* __restrict__ AT, const int* __restrict__ src2, const int* __restrict__ brow, int T0,
    const int* __restrict__ mask) {
  __shared__ T sB[1][KB][64];
  int b, w;
  xcd_matrix_map(b, w);
  if (mask && !mask[b]) return;
  gj_update_body<KB, 1, true, T>(X, Out, ostride, ldo, ocol0, Np, k0, AT, src2, brow, b, w, T0, T0 + 1, threadIdx.x, sB);
}

// One launch with two jobs on disjoint data: the first `npanel` workgroups run
// the panel chain of the next pair (one matrix each, as k_gj_pair_chain), the
// rest run column tiles [T0, T1) of the current pair's rank-2NB update, two
// 64-row bands per workgroup.  Workgroups dispatch in index order, so the
// latency-bound panel chains take their CUs first and the update fills the
// others -- the panel chain of pair q+1 hides behind update q.
template <int RPT, typename T>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_gj_fused(
    PairJob<T> P, int npanel, const T* __restrict__ X, T* __restrict__ Out, int Np, int k0,
    const T* __restrict__ AT, const int* __restrict__ src2, const int* __restrict__ brow, int T0, int T1,
    int ngroups) {
  if ((int)blockIdx.x < npanel) {
    if (P.mask && !P.mask[blockIdx.x]) return;
    gj_pair_chain_body<RPT, T>(P, blockIdx.x);
    return;
  }
  __shared__ T sB[2][2 * NB][64];
  const int u = blockIdx.x - npanel;
  const int band_pairs = ((Np + 63) / 64 + 1) / 2;
  const int per_b = band_pairs * ngroups;
  // Workgroups are dealt to the 8 XCDs round robin (npanel is a multiple of 8
  // here, so XCD = u % 8): the update workgroups of one matrix go to one XCD,
  // whose L2 then serves its staged B rows to all of its band pairs.
  int b, rem;
  if ((npanel & 7) == 0) {
    const int k = u >> 3;
    b = (u & 7) + 8 * (k / per_b);
    rem = k % per_b;
  } else {
    b = u / per_b;
    rem = u % per_b;
  }
  if (P.mask && !P.mask[b]) return;
  int jt0, jt1;
  group_tiles(T0, T1, rem % ngroups, ngroups, jt0, jt1);
  if (jt0 >= jt1) return;
  // a band past Np (odd band count) only stages B tiles and joins the barriers
  gj_update_body<2 * NB, 2, false, T>(X, Out, (long long)Np * Np, Np, 0, Np, k0, AT, src2, brow, b, 2 * (rem / ngroups), jt0,
                            jt1, threadIdx.x, sB);
}

// out[b][r][c] = in[b][r][pinv[c]], converted to the output type.
template <typename Tin, typename Tout>
__global__ void k_unpermute(const Tin* __restrict__ in, Tout* __restrict__ out, int Np,
                            const int* __restrict__ pinv, const int* __restrict__ mask) {
  const int b = blockIdx.y;
  if (mask && !mask[b]) return;
  const size_t n2 = (size_t)Np * Np;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n2; idx += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(idx / Np);
    const int c = (int)(idx % Np);
    out[b * n2 + idx] = (Tout)in[b * n2 + (size_t)r * Np + pinv[(size_t)b * Np + c]];
  }
}

// Live timing of the dominant kernel (k_gj_fused: a pair's update beside the
// next pair's panel chain) with HIP events: every `every`-th inversion has each
// of its fused launches bracketed by an event pair, recorded with the launch's
// algorithmic bytes and flop.
struct UpdateProfile {
  std::mutex mu;
  std::atomic<bool> on{false};
  int every = 1;
  long long calls = 0;
  std::vector<hipEvent_t> events;  // pairs: [2i] start, [2i+1] stop
  std::vector<double> bytes, flop;  // per timed launch
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
  g_prof.bytes.resize(g_prof.events.size() / 2);
  g_prof.flop.resize(g_prof.events.size() / 2);
  g_prof.on = true;
  return TICC_OK;
}

int profile_end(double* total_ms, double* union_ms, int* count, double* total_bytes, double* total_flop) {
  std::lock_guard<std::mutex> lock(g_prof.mu);
  g_prof.on = false;
  TICC_CHECK_HIP(hipDeviceSynchronize());
  double tot = 0.0, nbytes = 0.0, nflop = 0.0;
  // each launch as an interval on the device clock, relative to the first
  // timed launch's start event (events of every stream share that clock)
  std::vector<std::pair<double, double>> iv;
  iv.reserve(g_prof.used);
  for (size_t i = 0; i < g_prof.used; ++i) {
    nbytes += g_prof.bytes[i];
    nflop += g_prof.flop[i];
    float ms = 0.0f, t0 = 0.0f;
    TICC_CHECK_HIP(hipEventElapsedTime(&ms, g_prof.events[2 * i], g_prof.events[2 * i + 1]));
    TICC_CHECK_HIP(hipEventElapsedTime(&t0, g_prof.events[0], g_prof.events[2 * i]));
    tot += ms;
    iv.emplace_back((double)t0, (double)t0 + (double)ms);
  }
  // length of the union of the intervals: the time during which at least one
  // timed launch ran (equal to the sum when launches never overlap)
  std::sort(iv.begin(), iv.end());
  double uni = 0.0, lo = 0.0, hi = 0.0;
  bool open = false;
  for (const auto& x : iv) {
    if (!open || x.first > hi) {
      if (open) uni += hi - lo;
      lo = x.first;
      hi = x.second;
      open = true;
    } else if (x.second > hi) {
      hi = x.second;
    }
  }
  if (open) uni += hi - lo;
  *total_ms = tot;
  *union_ms = uni;
  *count = (int)g_prof.used;
  *total_bytes = nbytes;
  *total_flop = nflop;
  g_prof.used = 0;
  return TICC_OK;
}

namespace {
template <typename T>
struct PanelMaps {  // optional outputs of a panel launch
  T* PoutT = nullptr;
  long long tstride = 0;
  const int* prev_src = nullptr;
  const int* prev_piv = nullptr;
  int prev_k0 = 0;
  int* src2 = nullptr;
  int* brow = nullptr;
};

// Tile walk of the update launches: the workgroup count a launch's column tiles
// are spread over (fewer, longer walks amortise each band's prologue; more,
// shorter ones fill the chip at small batch).  Per calling thread, 0 = default.
constexpr int kWalkFused = 512;    // about two workgroups per CU beside the panel chains
constexpr int kWalkUpdate = 2048;  // standalone update launches
thread_local int t_walk_fused = 0;
thread_local int t_walk_update = 0;

template <typename T>
int launch_panel(const T* X, long long bstride, int ld, int col0, int batch, int Np, int k0, T* pout,
                 int* src, int* piv, int* perm, int* pinv, const PanelMaps<T>& m, const int* mask, hipStream_t stream) {
#define TICC_PANEL_ARGS                                                                                          \
  X, bstride, ld, col0, Np, k0, pout, src, piv, perm, pinv, m.PoutT, m.tstride, m.prev_src, m.prev_piv, m.prev_k0, \
      m.src2, m.brow, mask
  if (Np <= 256)
    hipLaunchKernelGGL((k_gj_panel<256, 1, T>), dim3(batch), dim3(256), 0, stream, TICC_PANEL_ARGS);
  else if (Np <= 512)
    hipLaunchKernelGGL((k_gj_panel<512, 1, T>), dim3(batch), dim3(512), 0, stream, TICC_PANEL_ARGS);
  else if (Np <= 1024)
    hipLaunchKernelGGL((k_gj_panel<512, 2, T>), dim3(batch), dim3(512), 0, stream, TICC_PANEL_ARGS);
  else
    hipLaunchKernelGGL((k_gj_panel<512, 3, T>), dim3(batch), dim3(512), 0, stream, TICC_PANEL_ARGS);
#undef TICC_PANEL_ARGS
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

template <int KB, typename T>
int launch_update(const T* cur, T* out, long long ostride, int ldo, int ocol0, int batch, int Np, int k0,
                  const T* AT, const int* src2, const int* brow, int T0, int T1, const int* mask,
                  hipStream_t stream) {
  // standalone launches run one band per workgroup (faster alone than the
  // two-band form, which pays off beside a panel in k_gj_fused)
  constexpr int NBANDS = 1;
  const int units = ((Np + 63) / 64 + NBANDS - 1) / NBANDS;  // band groups of NBANDS bands
  const int span = T1 - T0;
  if (span == 1) {
    hipLaunchKernelGGL((k_gj_update_one<KB, T>), dim3(units, batch), dim3(256), 0, stream, cur, out, ostride, ldo, ocol0,
                       Np, k0, AT, src2, brow, T0, mask);
    TICC_CHECK_LAUNCH();
    return TICC_OK;
  }
  // about walk / NBANDS workgroups: each walks span/ngroups tiles
  const int walk = t_walk_update > 0 ? t_walk_update : kWalkUpdate;
  int ngroups = (walk / NBANDS + units * batch - 1) / (units * batch);
  if (ngroups > span) ngroups = span;
  if (ngroups < 1) ngroups = 1;
  hipLaunchKernelGGL((k_gj_update<KB, NBANDS, T>), dim3(units * ngroups, batch), dim3(256 * NBANDS), 0, stream, cur, out,
                     ostride, ldo, ocol0, Np, k0, AT, src2, brow, T0, T1, ngroups, mask);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

template <typename T>
struct Carve {  // scratch carve-up of gj_invert_raw
  T *P1, *P2, *Z, *L, *AT[2];
  int *s1, *s2, *src2[2], *p1, *p2, *brow[2];
};

template <typename T>
Carve<T> carve(T* pout, int* src, int* piv, int batch, int Np) {
  Carve<T> c;
  const size_t ps = (size_t)batch * Np * NB;
  c.P1 = pout;
  c.P2 = pout + ps;
  c.Z = pout + 2 * ps;
  c.L = pout + 3 * ps;      // Np x 2NB per matrix: look-ahead columns of the next pair
  c.AT[0] = pout + 5 * ps;  // 2NB x Np per matrix, one set per pair parity
  c.AT[1] = pout + 7 * ps;
  const size_t ms = (size_t)batch * Np;
  c.s1 = src;
  c.s2 = src + ms;
  c.src2[0] = src + 2 * ms;
  c.src2[1] = src + 3 * ms;
  const size_t ls = (size_t)batch * NB;
  c.p1 = piv;
  c.p2 = piv + ls;
  c.brow[0] = piv + 2 * ls;  // 2NB per matrix
  c.brow[1] = piv + 4 * ls;
  return c;
}

template <typename T>
PairJob<T> pair_job(const T* Xc, long long xs, int ldx, int xcol, int Np, int k0, const Carve<T>& c, int par,
                    int* perm, int* pinv, const int* mask) {
  PairJob<T> P;
  P.mask = mask;
  P.Xc = Xc;
  P.xs = xs;
  P.ldx = ldx;
  P.xcol = xcol;
  P.Np = Np;
  P.k0 = k0;
  P.single = 0;
  P.P1 = c.P1;
  P.P2 = c.P2;
  P.Z = c.Z;
  P.AT = c.AT[par];
  P.s1 = c.s1;
  P.s2 = c.s2;
  P.p1 = c.p1;
  P.p2 = c.p2;
  P.src2 = c.src2[par];
  P.brow = c.brow[par];
  P.perm = perm;
  P.pinv = pinv;
  return P;
}

template <typename T>
int launch_pair_chain(const PairJob<T>& P, int batch, hipStream_t stream) {
  if (P.Np <= 512)
    hipLaunchKernelGGL((k_gj_pair_chain<1, T>), dim3(batch), dim3(512), 0, stream, P);
  else if (P.Np <= 1024)
    hipLaunchKernelGGL((k_gj_pair_chain<2, T>), dim3(batch), dim3(512), 0, stream, P);
  else
    hipLaunchKernelGGL((k_gj_pair_chain<3, T>), dim3(batch), dim3(512), 0, stream, P);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

template <typename T>
int single_step(T* cur, T* nxt, int batch, int Np, int k0, const Carve<T>& c, int* perm, int* pinv,
                const int* mask, hipStream_t st) {
  const long long n2 = (long long)Np * Np;
  PanelMaps<T> m;
  m.PoutT = c.AT[0];
  m.tstride = 2LL * NB * Np;
  m.src2 = c.src2[0];
  m.brow = c.brow[0];
  TICC_RETURN_IF(launch_panel(cur, n2, Np, k0, batch, Np, k0, c.P1, c.s1, c.p1, perm, pinv, m, mask, st));
  return launch_update<NB, T>(cur, nxt, n2, Np, 0, batch, Np, k0, c.AT[0], c.src2[0], c.brow[0], 0, (Np + 63) / 64,
                              mask, st);
}

// Fused launch: the panel chain P beside column tiles [T0, T1) of the rank-2NB
// update of the pair at k0.
// `slot` >= 0: the profile slot reserved for this launch (its events, bytes and
// flop entries), -1: not timed.
template <typename T>
int launch_fused(const PairJob<T>& P, const T* X, T* Out, int batch, int Np, int k0, const T* AT,
                 const int* src2, const int* brow, int T0, int T1, long long slot, hipStream_t stream) {
  const int band_pairs = ((Np + 63) / 64 + 1) / 2;
  // about two workgroups per CU: longer tile walks amortise each band's
  // prologue (A fragments, first C and B tiles) better than more, shorter ones
  const int walk = t_walk_fused > 0 ? t_walk_fused : kWalkFused;
  int ngroups = (walk + band_pairs * batch - 1) / (band_pairs * batch);
  if (ngroups > T1 - T0) ngroups = T1 - T0;
  if (ngroups < 1) ngroups = 1;
  const int grid = batch + band_pairs * ngroups * batch;
  if (slot >= 0) {
    // algorithmic work: the update reads X and writes the partner buffer on its
    // columns and reads the 2NB coefficient columns; the panel chain reads its
    // 2NB (or NB) look-ahead columns and writes 2NB coefficient columns
    const double cols = (double)((T1 * 64 < Np ? T1 * 64 : Np) - T0 * 64);
    const double kp = P.single ? NB : 2 * NB;
    g_prof.bytes[slot] = (double)sizeof(T) * batch * (2.0 * Np * cols + 2.0 * NB * Np + 2.0 * kp * Np);
    g_prof.flop[slot] = 2.0 * batch * Np * cols * 2 * NB;
    TICC_CHECK_HIP(hipEventRecord(g_prof.events[2 * slot], stream));
  }
  if (Np <= 512)
    hipLaunchKernelGGL((k_gj_fused<1, T>), dim3(grid), dim3(512), 0, stream, P, batch, X, Out, Np, k0, AT, src2, brow,
                       T0, T1, ngroups);
  else if (Np <= 1024)
    hipLaunchKernelGGL((k_gj_fused<2, T>), dim3(grid), dim3(512), 0, stream, P, batch, X, Out, Np, k0, AT, src2, brow,
                       T0, T1, ngroups);
  else
    hipLaunchKernelGGL((k_gj_fused<3, T>), dim3(grid), dim3(512), 0, stream, P, batch, X, Out, Np, k0, AT, src2, brow,
                       T0, T1, ngroups);
  TICC_CHECK_LAUNCH();
  if (slot >= 0) TICC_CHECK_HIP(hipEventRecord(g_prof.events[2 * slot + 1], stream));
  return TICC_OK;
}

template <typename T>
int gj_invert_t(T* bufA, T* bufB, int batch, int Np, T** result, int* perm, int* pinv, T* pout, int* src, int* piv,
                const int* mask, hipStream_t stream) {
  if (Np % NB != 0 || Np > MAX_NP || Np <= 0 || batch <= 0)
    return fail(TICC_ERR_VALUE, "gj_invert_raw: invalid Np=%d batch=%d", Np, batch);
  const Carve<T> c = carve<T>(pout, src, piv, batch, Np);
  const long long n2 = (long long)Np * Np;
  const int tiles = (Np + 63) / 64;
  T* buf[2] = {bufA, bufB};
  const int steps = Np / NB;
  const int pairs = steps / 2;
  // TICC_GJ_SINGLE_STEPS=1: every block step on its own -- its panel, then a
  // rank-NB update of the whole matrix -- with no pairing of steps and no
  // look-ahead.  The same elimination and pivots in another FP64 ordering (the
  // pair composition and the MFMA accumulation order change); slower, used to
  // measure the device's own rounding spread.
  if (const char* env = getenv("TICC_GJ_SINGLE_STEPS")) {
    if (env[0] == '1') {
      T* cur = bufA;
      T* nxt = bufB;
      for (int s = 0; s < steps; ++s) {
        TICC_RETURN_IF(single_step(cur, nxt, batch, Np, s * NB, c, perm, pinv, mask, stream));
        std::swap(cur, nxt);
      }
      *result = cur;
      return TICC_OK;
    }
  }
  // Profile slots of this inversion's fused launches (one per pair but the last,
  // plus the last pair's when a trailing single step follows), reserved under
  // the lock so that inversions on concurrent lanes never share a slot.
  const int n_fused = pairs > 0 ? pairs - 1 + (steps & 1) : 0;
  long long slot = -1;
  if (g_prof.on && n_fused > 0 && !mask) {  // masked (fallback) inversions are not timed
    std::lock_guard<std::mutex> lock(g_prof.mu);
    if (g_prof.on && (g_prof.calls++ % g_prof.every == 0) && g_prof.used + (size_t)n_fused <= g_prof.bytes.size()) {
      slot = (long long)g_prof.used;
      g_prof.used += (size_t)n_fused;
    }
  }
  auto next_slot = [&]() { return slot < 0 ? -1LL : slot++; };
  const long long ln = 2LL * NB * Np;  // look-ahead buffer: Np x 2NB per matrix
  if (pairs > 0) TICC_RETURN_IF(launch_pair_chain(pair_job<T>(buf[0], n2, Np, 0, Np, 0, c, 0, perm, pinv, mask), batch, stream));
  for (int q = 0; q < pairs; ++q) {
    const int k0 = 2 * q * NB;
    const int par = q & 1;
    T* cur = buf[q & 1];
    T* nxt = buf[(q + 1) & 1];
    if (q + 1 == pairs) {
      if (!(steps & 1)) {
        TICC_RETURN_IF(launch_update<2 * NB, T>(cur, nxt, n2, Np, 0, batch, Np, k0, c.AT[par], c.src2[par],
                                             c.brow[par], 0, tiles, mask, stream));
        break;
      }
      // trailing single step: its panel (on a look-ahead of its columns) runs
      // inside the last pair's update, then the rank-NB update
      const int ks = k0 + 2 * NB;
      TICC_RETURN_IF(launch_update<2 * NB, T>(cur, c.L, ln, 2 * NB, ks, batch, Np, k0, c.AT[par], c.src2[par],
                                           c.brow[par], ks / 64, ks / 64 + 1, mask, stream));
      PairJob<T> P = pair_job<T>(c.L, ln, 2 * NB, 0, Np, ks, c, par ^ 1, perm, pinv, mask);
      P.single = 1;
      TICC_RETURN_IF(
          launch_fused(P, cur, nxt, batch, Np, k0, c.AT[par], c.src2[par], c.brow[par], 0, tiles, next_slot(), stream));
      TICC_RETURN_IF(launch_update<NB, T>(nxt, cur, n2, Np, 0, batch, Np, ks, c.AT[par ^ 1], c.src2[par ^ 1],
                                       c.brow[par ^ 1], 0, tiles, mask, stream));
      *result = cur;
      return TICC_OK;
    }
    // The panel chain of pair q+1 (columns kn, kn + NB) runs inside update q:
    // its columns after pair q come from a one-tile look-ahead of update q.
    const int kn = k0 + 2 * NB;
    TICC_RETURN_IF(launch_update<2 * NB, T>(cur, c.L, ln, 2 * NB, kn, batch, Np, k0, c.AT[par], c.src2[par],
                                         c.brow[par], kn / 64, kn / 64 + 1, mask, stream));
    TICC_RETURN_IF(launch_fused(pair_job<T>(c.L, ln, 2 * NB, 0, Np, kn, c, par ^ 1, perm, pinv, mask), cur, nxt, batch, Np, k0,
                                c.AT[par], c.src2[par], c.brow[par], 0, tiles, next_slot(), stream));
  }
  T* cur = buf[pairs & 1];
  if (steps & 1) {  // trailing single step: rank-NB update
    TICC_RETURN_IF(single_step(cur, buf[(pairs + 1) & 1], batch, Np, (steps - 1) * NB, c, perm, pinv, mask, stream));
    cur = buf[(pairs + 1) & 1];
  }
  *result = cur;
  return TICC_OK;
}

}  // namespace

int gj_invert_raw(double* bufA, double* bufB, int batch, int Np, double** result, int* perm, int* pinv, double* pout,
                  int* src, int* piv, hipStream_t stream) {
  return gj_invert_t<double>(bufA, bufB, batch, Np, result, perm, pinv, pout, src, piv, nullptr, stream);
}

int gj_invert_raw_masked(double* bufA, double* bufB, int batch, int Np, double** result, int* perm, int* pinv,
                         double* pout, int* src, int* piv, const int* mask, hipStream_t stream) {
  return gj_invert_t<double>(bufA, bufB, batch, Np, result, perm, pinv, pout, src, piv, mask, stream);
}

int gj_invert_raw_f32(float* bufA, float* bufB, int batch, int Np, float** result, int* perm, int* pinv, float* pout,
                      int* src, int* piv, hipStream_t stream) {
  return gj_invert_t<float>(bufA, bufB, batch, Np, result, perm, pinv, pout, src, piv, nullptr, stream);
}

int gj_set_walk(int fused_target, int update_target) {
  if (fused_target < 0 || fused_target > 65536 || update_target < 0 || update_target > 65536)
    return fail(TICC_ERR_VALUE, "tile walk targets must lie in [0, 65536], got %d and %d", fused_target, update_target);
  t_walk_fused = fused_target;
  t_walk_update = update_target;
  return TICC_OK;
}

namespace {
template <typename Tin, typename Tout>
int unpermute_t(const Tin* in, Tout* out, int batch, int Np, const int* pinv, const int* mask, hipStream_t stream) {
  const long long n2 = (long long)Np * Np;
  int blocks = (int)((n2 + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL((k_unpermute<Tin, Tout>), dim3(blocks, batch), dim3(256), 0, stream, in, out, Np, pinv, mask);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}
}  // namespace

int gj_unpermute(const double* in, double* out, int batch, int Np, const int* pinv, hipStream_t stream) {
  return unpermute_t<double, double>(in, out, batch, Np, pinv, nullptr, stream);
}

int gj_unpermute_masked(const double* in, double* out, int batch, int Np, const int* pinv, const int* mask,
                        hipStream_t stream) {
  return unpermute_t<double, double>(in, out, batch, Np, pinv, mask, stream);
}

int gj_unpermute_f32(const float* in, double* out, int batch, int Np, const int* pinv, hipStream_t stream) {
  return unpermute_t<float, double>(in, out, batch, Np, pinv, nullptr, stream);
}

int gj_unpermute_f32f32(const float* in, float* out, int batch, int Np, const int* pinv, hipStream_t stream) {
  return unpermute_t<float, float>(in, out, batch, Np, pinv, nullptr, stream);
}

}  // namespace ticc
