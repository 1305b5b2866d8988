// Asymptotic matching, kernel C: BF->SF transform, reference functions,
// closed-channel elimination and K -> S for a batch of energies.
//
// Per energy e (channels ordered closed-first, each group ascending):
//   Ysf   = B^T Y B
//   Z     = (Y_cc - diag Gamma)^-1                         (batched GJ)
//   Yeff  = Y_oo - Y_oc Z Y_co
//   K     = -(Yeff diag N - diag N')^-1 (Yeff diag J - diag J')
//   S     = (I + K^2)^-1 (I - K^2) - 2i (I + K^2)^-1 K     (== (I+iK)^-1 (I-iK))
//   then two FP64 refinement sweeps against the residual of (I + iK) S = I - iK:
//     E = (I - iK) - (I + iK) S,   S += (I + K^2)^-1 (I - iK) E
//   (I + K^2 squares the conditioning of I + iK near a pole of K; the residual
//   form restores the accuracy of a direct complex solve).
// Ragged open/closed counts are handled by identity (inverted operands) or
// zero (multiplied operands) padding inside uniform Np x Np batches.
#include <cmath>
#include <vector>

#include "bessel.h"
#include "ticc_common.h"

namespace ticc {
namespace {

constexpr int ECHUNK = 32;

// J, N, J', N' (open) or Gamma (closed) for every (energy, channel).
__global__ void k_reference(int nE, int N, const double* __restrict__ E, const double* __restrict__ Eint,
                            const double* __restrict__ L, double two_mu, double Rm, double* __restrict__ Jv,
                            double* __restrict__ Nv, double* __restrict__ Jd, double* __restrict__ Nd,
                            double* __restrict__ G, int* __restrict__ flag) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nE * N) return;
  const int e = idx / N;
  const int i = idx % N;
  const double diff = E[e] - Eint[i];
  const double k = sqrt(two_mu * fabs(diff));
  const double x = k * Rm;
  Jv[idx] = Nv[idx] = Jd[idx] = Nd[idx] = G[idx] = 0.0;
  if (diff > 0.0) {
    const double nu = L[i] + 0.5;
    bessel::JY r = bessel::bessel_jy(nu, x);
    const double scale = sqrt(bessel::kPi * x / 2.0);
    const double jr = scale * r.j;
    const double yr = scale * r.y;
    const double jrp = scale * (r.jp + 0.5 * r.j / x);
    const double yrp = scale * (r.yp + 0.5 * r.y / x);
    const double sk = sqrt(k);
    Jv[idx] = jr / sk;
    Nv[idx] = yr / sk;
    Jd[idx] = sk * jrp;
    Nd[idx] = sk * yrp;
    if (!r.ok || !isfinite(jr) || !isfinite(yr) || !isfinite(jrp) || !isfinite(yrp)) atomicOr(flag, 1);
  } else if (diff < 0.0) {
    int ok = 1;
    const double kl = bessel::bessel_k_logd(L[i] + 0.5, x, &ok);
    G[idx] = 0.5 / Rm + k * kl;
    if (!ok || !isfinite(kl)) atomicOr(flag, 1);
  }
}

__global__ void k_bessel_eval(int n, const double* __restrict__ ell, const double* __restrict__ x,
                              double* __restrict__ out, const double* __restrict__ nu, double* __restrict__ kout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (out) {
    const double xv = x[i];
    bessel::JY r = bessel::bessel_jy(ell[i] + 0.5, xv);
    const double scale = sqrt(bessel::kPi * xv / 2.0);
    out[4 * i + 0] = scale * r.j;
    out[4 * i + 1] = scale * r.y;
    out[4 * i + 2] = scale * (r.jp + 0.5 * r.j / xv);
    out[4 * i + 3] = scale * (r.yp + 0.5 * r.y / xv);
  }
  if (kout) {
    int ok = 1;
    kout[i] = bessel::bessel_k_logd(nu[i], x[i], &ok);
  }
}

// Split Ysf into the padded closed operator, Y_co, Y_oc and Y_oo.
__global__ void k_build_blocks(const double* __restrict__ Ysf, const int* __restrict__ order,
                               const int* __restrict__ ncl, const double* __restrict__ G, int N, int Np,
                               double* __restrict__ Mc, double* __restrict__ Yco, double* __restrict__ Yoc,
                               double* __restrict__ Yoo) {
  const int e = blockIdx.y;
  const int nc = ncl[e];
  const int no = N - nc;
  const size_t n2 = (size_t)Np * Np;
  const int* ord = order + (size_t)e * N;
  const double* Y = Ysf + e * n2;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n2; idx += (size_t)gridDim.x * blockDim.x) {
    const int a = (int)(idx / Np);
    const int b = (int)(idx % Np);
    const double d = (a == b) ? 1.0 : 0.0;
    double mc = d, co = 0.0, oc = 0.0, oo = 0.0;
    if (a < nc && b < nc) {
      mc = Y[(size_t)ord[a] * Np + ord[b]];
      if (a == b) mc -= G[(size_t)e * N + ord[a]];
    }
    if (a < nc && b < no) co = Y[(size_t)ord[a] * Np + ord[nc + b]];
    if (a < no && b < nc) oc = Y[(size_t)ord[nc + a] * Np + ord[b]];
    if (a < no && b < no) oo = Y[(size_t)ord[nc + a] * Np + ord[nc + b]];
    Mc[e * n2 + idx] = mc;
    Yco[e * n2 + idx] = co;
    Yoc[e * n2 + idx] = oc;
    Yoo[e * n2 + idx] = oo;
  }
}

// L = Yeff diag N - diag N' (identity padded), R = Yeff diag J - diag J' (zero padded).
__global__ void k_build_lr(const double* __restrict__ Yeff, const int* __restrict__ order,
                           const int* __restrict__ ncl, const double* __restrict__ Jv, const double* __restrict__ Nv,
                           const double* __restrict__ Jd, const double* __restrict__ Nd, int N, int Np,
                           double* __restrict__ Lm, double* __restrict__ Rm) {
  const int e = blockIdx.y;
  const int nc = ncl[e];
  const int no = N - nc;
  const size_t n2 = (size_t)Np * Np;
  const int* ord = order + (size_t)e * N + nc;
  const size_t eo = (size_t)e * N;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n2; idx += (size_t)gridDim.x * blockDim.x) {
    const int a = (int)(idx / Np);
    const int b = (int)(idx % Np);
    double l = (a == b) ? 1.0 : 0.0, r = 0.0;
    if (a < no && b < no) {
      const double y = Yeff[e * n2 + idx];
      l = y * Nv[eo + ord[b]];
      r = y * Jv[eo + ord[b]];
      if (a == b) {
        l -= Nd[eo + ord[a]];
        r -= Jd[eo + ord[a]];
      }
    }
    Lm[e * n2 + idx] = l;
    Rm[e * n2 + idx] = r;
  }
}

// M = I + K^2 (identity padded), P = I - K^2 (zero padded).
__global__ void k_build_mp(const double* __restrict__ KK, const int* __restrict__ ncl, int N, int Np,
                           double* __restrict__ Mm, double* __restrict__ Pm) {
  const int e = blockIdx.y;
  const int no = N - ncl[e];
  const size_t n2 = (size_t)Np * Np;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n2; idx += (size_t)gridDim.x * blockDim.x) {
    const int a = (int)(idx / Np);
    const int b = (int)(idx % Np);
    const double d = (a == b) ? 1.0 : 0.0;
    double m = d, p = 0.0;
    if (a < no && b < no) {
      const double kk = KK[e * n2 + idx];
      m = d + kk;
      p = d - kk;
    }
    Mm[e * n2 + idx] = m;
    Pm[e * n2 + idx] = p;
  }
}

// Residual of (I + iK) S = I - iK on the open block, from KSi = K Im S and
// KSr = K Re S (both overwritten): Re E = I - Re S + K Im S,
// Im E = -K - Im S - K Re S.  Zero outside the open block.
__global__ void k_smat_residual(const double* __restrict__ Sr, const double* __restrict__ Si,
                                const double* __restrict__ K, const int* __restrict__ ncl, int N, int Np,
                                double* __restrict__ KSi_Er, double* __restrict__ KSr_Ei) {
  const int e = blockIdx.y;
  const int no = N - ncl[e];
  const size_t n2 = (size_t)Np * Np;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n2; idx += (size_t)gridDim.x * blockDim.x) {
    const int a = (int)(idx / Np);
    const int b = (int)(idx % Np);
    const size_t o = e * n2 + idx;
    double er = 0.0, ei = 0.0;
    if (a < no && b < no) {
      er = ((a == b) ? 1.0 : 0.0) - Sr[o] + KSi_Er[o];
      ei = -K[o] - Si[o] - KSr_Ei[o];
    }
    KSi_Er[o] = er;
    KSr_Ei[o] = ei;
  }
}

int grid1(long long n) {
  long long b = (n + 255) / 256;
  if (b > 2048) b = 2048;
  return (int)(b < 1 ? 1 : b);
}

}  // namespace

int bessel_eval(int n, const double* ell, const double* x, double* out, const double* nu, double* kout,
                hipStream_t stream) {
  if (n <= 0) return TICC_OK;
  hipLaunchKernelGGL(k_bessel_eval, dim3((n + 127) / 128), dim3(128), 0, stream, n, ell, x, out, nu, kout);
  TICC_CHECK_LAUNCH();
  return TICC_OK;
}

int get_smat(int nE, int N, const double* Y, const double* B, double Rmatch, const double* E_host, double mu,
             const double* Eint_host, const double* L_host, double* S_re, double* S_im, int* n_open,
             hipStream_t stream) {
  const int Np = round_up(N, GJ_NB);
  if (Np > MAX_NP) return fail(TICC_ERR_UNSUPPORTED, "n_channel=%d exceeds the supported maximum %d", N, MAX_NP);
  // host-side partition (smatrix.py:50-57): closed first, then open, each ascending.
  std::vector<int> order((size_t)nE * N), ncl(nE);
  for (int e = 0; e < nE; ++e) {
    int p = 0;
    for (int i = 0; i < N; ++i) {
      const double diff = E_host[e] - Eint_host[i];
      if (diff == 0.0)
        return fail(TICC_ERR_VALUE, "Asymptotic matching is undefined exactly at a channel threshold Etot=%.17g",
                    E_host[e]);
      if (diff < 0.0) order[(size_t)e * N + p++] = i;
    }
    ncl[e] = p;
    for (int i = 0; i < N; ++i)
      if (E_host[e] - Eint_host[i] > 0.0) order[(size_t)e * N + p++] = i;
    n_open[e] = N - ncl[e];
  }
  const int EC = nE < ECHUNK ? nE : ECHUNK;
  const size_t n2 = (size_t)Np * Np;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  size_t ob[7];
  for (int i = 0; i < 7; ++i) ob[i] = take(sizeof(double) * EC * n2);
  const size_t oB = take(sizeof(double) * n2);
  const size_t oRef = take(sizeof(double) * 5 * (size_t)nE * N);
  const size_t oE = take(sizeof(double) * nE);
  const size_t oEint = take(sizeof(double) * N);
  const size_t oL = take(sizeof(double) * N);
  const size_t oOrd = take(sizeof(int) * (size_t)nE * N);
  const size_t oNc = take(sizeof(int) * nE);
  const size_t oP = take(sizeof(double) * GJ_POUT_PANELS * EC * Np * GJ_NB);
  const size_t oSrc = take(sizeof(int) * GJ_SRC_MAPS * EC * Np);
  const size_t oPiv = take(sizeof(int) * GJ_PIV_LISTS * EC * GJ_NB);
  const size_t oPerm = take(sizeof(int) * EC * Np);
  const size_t oPinv = take(sizeof(int) * EC * Np);
  const size_t oFlag = take(sizeof(int));
  void* base = nullptr;
  TICC_RETURN_IF(workspace(WS_MATCH, off, &base));
  char* w = static_cast<char*>(base);
  double* b[7];
  for (int i = 0; i < 7; ++i) b[i] = reinterpret_cast<double*>(w + ob[i]);
  double* Bp = reinterpret_cast<double*>(w + oB);
  double* ref = reinterpret_cast<double*>(w + oRef);
  double *Jv = ref, *Nv = ref + (size_t)nE * N, *Jd = ref + 2 * (size_t)nE * N, *Nd = ref + 3 * (size_t)nE * N,
         *G = ref + 4 * (size_t)nE * N;
  double* dE = reinterpret_cast<double*>(w + oE);
  double* dEint = reinterpret_cast<double*>(w + oEint);
  double* dL = reinterpret_cast<double*>(w + oL);
  int* dOrd = reinterpret_cast<int*>(w + oOrd);
  int* dNc = reinterpret_cast<int*>(w + oNc);
  double* Pout = reinterpret_cast<double*>(w + oP);
  int* src = reinterpret_cast<int*>(w + oSrc);
  int* piv = reinterpret_cast<int*>(w + oPiv);
  int* perm = reinterpret_cast<int*>(w + oPerm);
  int* pinv = reinterpret_cast<int*>(w + oPinv);
  int* flag = reinterpret_cast<int*>(w + oFlag);

  TICC_CHECK_HIP(hipMemcpyAsync(dE, E_host, sizeof(double) * nE, hipMemcpyHostToDevice, stream));
  TICC_CHECK_HIP(hipMemcpyAsync(dEint, Eint_host, sizeof(double) * N, hipMemcpyHostToDevice, stream));
  TICC_CHECK_HIP(hipMemcpyAsync(dL, L_host, sizeof(double) * N, hipMemcpyHostToDevice, stream));
  TICC_CHECK_HIP(hipMemcpyAsync(dOrd, order.data(), sizeof(int) * (size_t)nE * N, hipMemcpyHostToDevice, stream));
  TICC_CHECK_HIP(hipMemcpyAsync(dNc, ncl.data(), sizeof(int) * nE, hipMemcpyHostToDevice, stream));
  TICC_CHECK_HIP(hipMemsetAsync(flag, 0, sizeof(int), stream));
  hipLaunchKernelGGL(k_reference, dim3((nE * N + 127) / 128), dim3(128), 0, stream, nE, N, dE, dEint, dL, 2.0 * mu,
                     Rmatch, Jv, Nv, Jd, Nd, G, flag);
  TICC_CHECK_LAUNCH();
  if (B) TICC_RETURN_IF(pad_identity(B, 0, N, Bp, Np, 1, stream));

  const int ge = grid1((long long)n2);
  for (int e0 = 0; e0 < nE; e0 += EC) {
    const int ne = (nE - e0) < EC ? (nE - e0) : EC;
    const int* ordc = dOrd + (size_t)e0 * N;
    const int* ncc = dNc + e0;
    // 1. Ysf = B^T Y B  -> b[2]
    TICC_RETURN_IF(pad_identity(Y + (size_t)e0 * N * N, (long long)N * N, N, b[0], Np, ne, stream));
    if (B) {
      TICC_RETURN_IF(dgemm_batched(false, false, Np, Np, Np, 1.0, b[0], Np, n2, Bp, Np, 0, 0.0, b[1], Np, n2, ne,
                                   stream));
      TICC_RETURN_IF(dgemm_batched(true, false, Np, Np, Np, 1.0, Bp, Np, 0, b[1], Np, n2, 0.0, b[2], Np, n2, ne,
                                   stream));
    } else {
      TICC_CHECK_HIP(hipMemcpyAsync(b[2], b[0], sizeof(double) * ne * n2, hipMemcpyDeviceToDevice, stream));
    }
    // 2. blocks: Mc -> b0, Yco -> b1, Yoc -> b3, Yoo -> b4
    hipLaunchKernelGGL(k_build_blocks, dim3(ge, ne), dim3(256), 0, stream, b[2], ordc, ncc, G + (size_t)e0 * N, N, Np,
                       b[0], b[1], b[3], b[4]);
    TICC_CHECK_LAUNCH();
    // 3. Z = Mc^-1 -> b2
    double* res = nullptr;
    TICC_RETURN_IF(gj_invert_raw(b[0], b[5], ne, Np, &res, perm, pinv, Pout, src, piv, stream));
    TICC_RETURN_IF(gj_unpermute(res, b[2], ne, Np, pinv, stream));
    // 4. Yeff = Yoo - Yoc (Z Yco)
    TICC_RETURN_IF(dgemm_batched(false, false, Np, Np, Np, 1.0, b[2], Np, n2, b[1], Np, n2, 0.0, b[0], Np, n2, ne,
                                 stream));
    TICC_RETURN_IF(dgemm_batched(false, false, Np, Np, Np, -1.0, b[3], Np, n2, b[0], Np, n2, 1.0, b[4], Np, n2, ne,
                                 stream));
    // 5. L -> b1, R -> b2
    hipLaunchKernelGGL(k_build_lr, dim3(ge, ne), dim3(256), 0, stream, b[4], ordc, ncc, Jv + (size_t)e0 * N,
                       Nv + (size_t)e0 * N, Jd + (size_t)e0 * N, Nd + (size_t)e0 * N, N, Np, b[1], b[2]);
    TICC_CHECK_LAUNCH();
    // 6. Li -> b3 ; K = -Li R -> b4
    TICC_RETURN_IF(gj_invert_raw(b[1], b[5], ne, Np, &res, perm, pinv, Pout, src, piv, stream));
    TICC_RETURN_IF(gj_unpermute(res, b[3], ne, Np, pinv, stream));
    TICC_RETURN_IF(dgemm_batched(false, false, Np, Np, Np, -1.0, b[3], Np, n2, b[2], Np, n2, 0.0, b[4], Np, n2, ne,
                                 stream));
    // 7. KK -> b0 ; M -> b1, P -> b2
    TICC_RETURN_IF(dgemm_batched(false, false, Np, Np, Np, 1.0, b[4], Np, n2, b[4], Np, n2, 0.0, b[0], Np, n2, ne,
                                 stream));
    hipLaunchKernelGGL(k_build_mp, dim3(ge, ne), dim3(256), 0, stream, b[0], ncc, N, Np, b[1], b[2]);
    TICC_CHECK_LAUNCH();
    // 8. Mi -> b3 ; Re S = Mi P -> b0 ; Im S = -2 Mi K -> b5
    TICC_RETURN_IF(gj_invert_raw(b[1], b[5], ne, Np, &res, perm, pinv, Pout, src, piv, stream));
    TICC_RETURN_IF(gj_unpermute(res, b[3], ne, Np, pinv, stream));
    TICC_RETURN_IF(dgemm_batched(false, false, Np, Np, Np, 1.0, b[3], Np, n2, b[2], Np, n2, 0.0, b[0], Np, n2, ne,
                                 stream));
    TICC_RETURN_IF(dgemm_batched(false, false, Np, Np, Np, -2.0, b[3], Np, n2, b[4], Np, n2, 0.0, b[5], Np, n2, ne,
                                 stream));
    // 9. refinement: E -> (b1, b2); T = (I - iK) E -> (b1, b6); S += Mi T
    for (int it = 0; it < 2; ++it) {
      TICC_RETURN_IF(dgemm_batched(false, false, Np, Np, Np, 1.0, b[4], Np, n2, b[5], Np, n2, 0.0, b[1], Np, n2, ne,
                                   stream));
      TICC_RETURN_IF(dgemm_batched(false, false, Np, Np, Np, 1.0, b[4], Np, n2, b[0], Np, n2, 0.0, b[2], Np, n2, ne,
                                   stream));
      hipLaunchKernelGGL(k_smat_residual, dim3(ge, ne), dim3(256), 0, stream, b[0], b[5], b[4], ncc, N, Np, b[1],
                         b[2]);
      TICC_CHECK_LAUNCH();
      TICC_CHECK_HIP(hipMemcpyAsync(b[6], b[2], sizeof(double) * ne * n2, hipMemcpyDeviceToDevice, stream));
      TICC_RETURN_IF(dgemm_batched(false, false, Np, Np, Np, -1.0, b[4], Np, n2, b[1], Np, n2, 1.0, b[6], Np, n2, ne,
                                   stream));
      TICC_RETURN_IF(dgemm_batched(false, false, Np, Np, Np, 1.0, b[4], Np, n2, b[2], Np, n2, 1.0, b[1], Np, n2, ne,
                                   stream));
      TICC_RETURN_IF(dgemm_batched(false, false, Np, Np, Np, 1.0, b[3], Np, n2, b[1], Np, n2, 1.0, b[0], Np, n2, ne,
                                   stream));
      TICC_RETURN_IF(dgemm_batched(false, false, Np, Np, Np, 1.0, b[3], Np, n2, b[6], Np, n2, 1.0, b[5], Np, n2, ne,
                                   stream));
    }
    TICC_RETURN_IF(unpad(b[0], Np, S_re + (size_t)e0 * N * N, (long long)N * N, N, ne, stream));
    TICC_RETURN_IF(unpad(b[5], Np, S_im + (size_t)e0 * N * N, (long long)N * N, N, ne, stream));
  }
  TICC_RETURN_IF(check_finite(S_re, (long long)nE * N * N, flag, stream));
  TICC_RETURN_IF(check_finite(S_im, (long long)nE * N * N, flag, stream));
  int hflag = 0;
  TICC_CHECK_HIP(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, stream));
  TICC_CHECK_HIP(hipStreamSynchronize(stream));
  if (hflag) return fail(TICC_ERR_NONFINITE, "asymptotic matching produced non-finite values");
  return TICC_OK;
}

}  // namespace ticc
