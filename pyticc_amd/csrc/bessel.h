// FP64 Bessel functions of real order for the asymptotic matching (device code).
//
// Written from the published methods, not from any library source:
//  * J'_nu/J_nu by the continued fraction CF1 (the ratio J_{nu+1}/J_nu,
//    A&S 9.1.73) evaluated with the modified Lentz algorithm
//    (Thompson & Barnett, J. Comput. Phys. 64 (1986) 490);
//  * Y_mu, Y_{mu+1} for |mu| <= 1/2, x < 2 by Temme's power series
//    (N. M. Temme, J. Comput. Phys. 21 (1976) 343), with 1/Gamma(1 +- mu)
//    from the Taylor series of 1/Gamma (A&S 6.1.34 coefficients);
//  * for x >= 2, Steed's complex continued fraction CF2 for
//    (J' + iY')/(J + iY) (Barnett, Feng, Steed & Goldfarb, Comput. Phys.
//    Commun. 8 (1974) 377) plus the Wronskian J Y' - J' Y = 2/(pi x);
//  * K_{mu+1}/K_mu from Temme's series (x < 2) or Steed's CF2 for K
//    (Thompson & Barnett 1987), then the upward ratio recurrence
//    r_{k+1} = 2(mu+k+1)/x + 1/r_k (stable for K).
#pragma once

#include <hip/hip_runtime.h>

namespace ticc {
namespace bessel {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEps = 1.0e-16;
constexpr double kTiny = 1.0e-300;
constexpr int kMaxIt = 200000;

// Taylor coefficients c_k of 1/Gamma(z) = sum_{k>=1} c_k z^k (A&S 6.1.34).
__host__ __device__ inline double inv_gamma_coef(int k) {
  switch (k) {
    case 1: return 1.0;
    case 2: return 0.5772156649015329;
    case 3: return -0.6558780715202538;
    case 4: return -0.0420026350340952;
    case 5: return 0.1665386113822915;
    case 6: return -0.0421977345555443;
    case 7: return -0.0096219715278770;
    case 8: return 0.0072189432466630;
    case 9: return -0.0011651675918591;
    case 10: return -0.0002152416741149;
    case 11: return 0.0001280502823882;
    case 12: return -0.0000201348547807;
    case 13: return -0.0000012504934821;
    case 14: return 0.0000011330272320;
    case 15: return -0.0000002056338417;
    case 16: return 0.0000000061160950;
    case 17: return 0.0000000050020075;
    case 18: return -0.0000000011812746;
    case 19: return 0.0000000001043427;
    case 20: return 0.0000000000077823;
    case 21: return -0.0000000000036968;
    case 22: return 0.0000000000005100;
    case 23: return -0.0000000000000206;
    case 24: return -0.0000000000000054;
    case 25: return 0.0000000000000014;
    case 26: return 0.0000000000000001;
    default: return 0.0;
  }
}

// For |mu| <= 1/2: gp = 1/Gamma(1+mu), gm = 1/Gamma(1-mu),
// g1 = (gm - gp)/(2 mu) (limit at mu = 0), g2 = (gm + gp)/2.
__host__ __device__ inline void gamma_pieces(double mu, double* g1, double* g2, double* gp, double* gm) {
  // 1/Gamma(1+z) = sum_{k>=1} c_k z^{k-1}
  double even = 0.0, odd = 0.0;  // even: sum over odd k (even powers), odd: even k (odd powers)/mu
  const double m2 = mu * mu;
  // Horner in mu^2 over the two parities.
  for (int k = 25; k >= 1; k -= 2) even = even * m2 + inv_gamma_coef(k);
  for (int k = 26; k >= 2; k -= 2) odd = odd * m2 + inv_gamma_coef(k);
  // 1/Gamma(1+mu) = even + mu*odd ; 1/Gamma(1-mu) = even - mu*odd
  *gp = even + mu * odd;
  *gm = even - mu * odd;
  *g1 = -odd;
  *g2 = even;
}

struct JY {
  double j, y, jp, yp;  // J_nu, Y_nu, J'_nu, Y'_nu
  int ok;
};

// Hankel's asymptotic expansion (A&S 9.2.5, 9.2.9-10) for x >> nu^2:
// J = sqrt(2/(pi x)) (P cos chi - Q sin chi), Y = sqrt(2/(pi x)) (P sin chi + Q cos chi),
// chi = x - (nu/2 + 1/4) pi.  Terms a_k = prod_{j<=k} (4nu^2 - (2j-1)^2) / (k! (8x)^k).
__host__ __device__ inline void hankel_jy(double nu, double x, double* J, double* Y) {
  const double m = 4.0 * nu * nu;
  const double z8 = 8.0 * x;
  double P = 1.0, Q = 0.0, term = 1.0;
  for (int k = 1; k < 200; ++k) {
    term *= (m - (2.0 * k - 1.0) * (2.0 * k - 1.0)) / (k * z8);
    // k odd -> Q with sign (-1)^((k-1)/2); k even -> P with sign (-1)^(k/2)
    if (k & 1)
      Q += (((k - 1) / 2) % 2 ? -term : term);
    else
      P += ((k / 2) % 2 ? -term : term);
    if (fabs(term) < 1.0e-17 * (fabs(P) + fabs(Q))) break;
  }
  const double a = (0.5 * nu + 0.25) * kPi;
  const double cx = cos(x), sx = sin(x), ca = cos(a), sa = sin(a);
  const double cchi = cx * ca + sx * sa;  // cos(x - a)
  const double schi = sx * ca - cx * sa;  // sin(x - a)
  const double amp = sqrt(2.0 / (kPi * x));
  *J = amp * (P * cchi - Q * schi);
  *Y = amp * (P * schi + Q * cchi);
}

// Ordinary Bessel J_nu, Y_nu and derivatives, nu >= 0, x > 0.
__host__ __device__ inline JY bessel_jy(double nu, double x) {
  JY out{0.0, 0.0, 0.0, 0.0, 1};
  if (x >= 30.0 && nu < 0.5 * x) {
    // Oscillatory region: Hankel's expansion at the low orders mu, mu+1
    // (mu in [0,1), accurate for x >= 30), then the three-term recurrence
    // upward for both J and Y, which is stable while nu < x.
    const int nup = (x > (nu + 1.0) * (nu + 1.0)) ? 0 : (int)nu;
    const double mu = nu - nup;
    double jm, ym, j1, y1;
    hankel_jy(mu, x, &jm, &ym);
    hankel_jy(mu + 1.0, x, &j1, &y1);
    for (int i = 1; i <= nup; ++i) {
      const double c = 2.0 * (mu + i) / x;
      const double jt = c * j1 - jm;
      const double yt = c * y1 - ym;
      jm = j1;
      ym = y1;
      j1 = jt;
      y1 = yt;
    }
    out.j = jm;
    out.y = ym;
    out.jp = nu / x * jm - j1;
    out.yp = nu / x * ym - y1;
    return out;
  }
  const double xi = 1.0 / x;
  const double xi2 = 2.0 * xi;
  const double w = xi2 / kPi;  // Wronskian 2/(pi x)
  int nl;
  if (x < 2.0)
    nl = (int)(nu + 0.5);
  else {
    nl = (int)(nu - x + 1.5);
    if (nl < 0) nl = 0;
  }
  const double mu = nu - nl;
  // CF1 for f = J'_nu / J_nu (modified Lentz), sign of J_nu tracked in isign.
  int isign = 1;
  double h = nu * xi;
  if (h < kTiny) h = kTiny;
  double b = xi2 * nu;
  double d = 0.0;
  double c = h;
  int it = 0;
  for (it = 1; it <= kMaxIt; ++it) {
    b += xi2;
    d = b - d;
    if (fabs(d) < kTiny) d = kTiny;
    c = b - 1.0 / c;
    if (fabs(c) < kTiny) c = kTiny;
    d = 1.0 / d;
    const double del = c * d;
    h = del * h;
    if (d < 0.0) isign = -isign;
    if (fabs(del - 1.0) < kEps) break;
  }
  if (it > kMaxIt) out.ok = 0;
  // Downward recurrence from nu to mu with unnormalised values.
  double jl = isign * 1.0e-280;
  double jpl = h * jl;
  const double jl_top = jl;
  double scale_log = 0.0;  // accumulated rescaling factor exponent (base 1e250)
  double fct = nu * xi;
  for (int l = nl; l >= 1; --l) {
    const double jtmp = fct * jl + jpl;
    fct -= xi;
    jpl = fct * jtmp - jl;
    jl = jtmp;
    if (fabs(jl) > 1.0e250) {
      jl *= 1.0e-250;
      jpl *= 1.0e-250;
      scale_log += 1.0;
    }
  }
  if (jl == 0.0) jl = kEps;
  const double f = jpl / jl;  // J'_mu / J_mu
  double rjmu, rymu, rymup, ry1;
  if (x < 2.0) {
    // Temme's series for Y_mu, Y_{mu+1}.
    const double x2 = 0.5 * x;
    const double pimu = kPi * mu;
    const double fact = (fabs(pimu) < kEps) ? 1.0 : pimu / sin(pimu);
    double dd = -log(x2);
    double e = mu * dd;
    const double fact2 = (fabs(e) < kEps) ? 1.0 : sinh(e) / e;
    double g1, g2, gp, gm;
    gamma_pieces(mu, &g1, &g2, &gp, &gm);
    double ff = 2.0 / kPi * fact * (g1 * cosh(e) + g2 * fact2 * dd);
    e = exp(e);
    double p = e / (gp * kPi);
    double q = 1.0 / (e * kPi * gm);
    const double pimu2 = 0.5 * pimu;
    const double fact3 = (fabs(pimu2) < kEps) ? 1.0 : sin(pimu2) / pimu2;
    const double r = kPi * pimu2 * fact3 * fact3;
    double cc = 1.0;
    dd = -x2 * x2;
    double sum = ff + r * q;
    double sum1 = p;
    int i;
    for (i = 1; i <= kMaxIt; ++i) {
      ff = (i * ff + p + q) / (i * (double)i - mu * mu);
      cc *= dd / i;
      p /= (i - mu);
      q /= (i + mu);
      const double del = cc * (ff + r * q);
      sum += del;
      const double del1 = cc * p - i * del;
      sum1 += del1;
      if (fabs(del) < (1.0 + fabs(sum)) * kEps) break;
    }
    if (i > kMaxIt) out.ok = 0;
    rymu = -sum;
    ry1 = -sum1 * xi2;
    rymup = mu * xi * rymu - ry1;
    rjmu = w / (rymup - f * rymu);
  } else {
    // Steed's CF2: p + i q = (J'_mu + i Y'_mu) / (J_mu + i Y_mu).
    double a = 0.25 - mu * mu;
    double p = -0.5 * xi;
    double q = 1.0;
    const double br = 2.0 * x;
    double bi = 2.0;
    double fct2 = a * xi / (p * p + q * q);
    double cr = br + q * fct2;
    double ci = bi + p * fct2;
    double den = br * br + bi * bi;
    double dr = br / den;
    double di = -bi / den;
    double dlr = cr * dr - ci * di;
    double dli = cr * di + ci * dr;
    double tmp = p * dlr - q * dli;
    q = p * dli + q * dlr;
    p = tmp;
    int i;
    for (i = 2; i <= kMaxIt; ++i) {
      a += 2 * (i - 1);
      bi += 2.0;
      dr = a * dr + br;
      di = a * di + bi;
      if (fabs(dr) + fabs(di) < kTiny) dr = kTiny;
      fct2 = a / (cr * cr + ci * ci);
      cr = br + cr * fct2;
      ci = bi - ci * fct2;
      if (fabs(cr) + fabs(ci) < kTiny) cr = kTiny;
      den = dr * dr + di * di;
      dr /= den;
      di /= -den;
      dlr = cr * dr - ci * di;
      dli = cr * di + ci * dr;
      tmp = p * dlr - q * dli;
      q = p * dli + q * dlr;
      p = tmp;
      if (fabs(dlr - 1.0) + fabs(dli) < kEps) break;
    }
    if (i > kMaxIt) out.ok = 0;
    const double gam = (p - f) / q;
    rjmu = sqrt(w / ((p - f) * gam + q));
    rjmu = (jl < 0.0) ? -rjmu : rjmu;
    rymu = rjmu * gam;
    rymup = rymu * (p + q / gam);
    ry1 = mu * xi * rymu - rymup;
  }
  const double fact_n = rjmu / jl;  // normalisation of the downward J values
  // J_nu = fact_n * jl_top * 1e250^scale_log
  double jn = jl_top * fact_n;
  for (double s = 0; s < scale_log; s += 1.0) jn *= 1.0e-250;
  out.j = jn;
  out.jp = h * jn;
  // Upward recurrence for Y from mu to nu.
  double ym = rymu, yp1 = ry1;
  for (int i = 1; i <= nl; ++i) {
    const double ytmp = (mu + i) * xi2 * yp1 - ym;
    ym = yp1;
    yp1 = ytmp;
  }
  out.y = ym;
  out.yp = nu * xi * ym - yp1;
  return out;
}

// K'_nu(x)/K_nu(x) for nu >= 0, x > 0.
__host__ __device__ inline double bessel_k_logd(double nu, double x, int* ok) {
  const int nl = (int)(nu + 0.5);
  const double mu = nu - nl;
  const double xi = 1.0 / x;
  double ratio;  // K_{mu+1} / K_mu
  if (x < 2.0) {
    const double x2 = 0.5 * x;
    const double pimu = kPi * mu;
    const double fact = (fabs(pimu) < kEps) ? 1.0 : pimu / sin(pimu);
    double dd = -log(x2);
    double e = mu * dd;
    const double fact2 = (fabs(e) < kEps) ? 1.0 : sinh(e) / e;
    double g1, g2, gp, gm;
    gamma_pieces(mu, &g1, &g2, &gp, &gm);
    double ff = fact * (g1 * cosh(e) + g2 * fact2 * dd);
    double sum = ff;
    e = exp(e);
    double p = 0.5 * e / gp;
    double q = 0.5 / (e * gm);
    double cc = 1.0;
    dd = x2 * x2;
    double sum1 = p;
    int i;
    for (i = 1; i <= kMaxIt; ++i) {
      ff = (i * ff + p + q) / (i * (double)i - mu * mu);
      cc *= dd / i;
      p /= (i - mu);
      q /= (i + mu);
      const double del = cc * ff;
      sum += del;
      const double del1 = cc * (p - i * ff);
      sum1 += del1;
      if (fabs(del) < fabs(sum) * kEps) break;
    }
    if (i > kMaxIt) *ok = 0;
    ratio = sum1 * 2.0 * xi / sum;
  } else {
    // Steed's CF2 for K: only the continued-fraction part h is needed for the ratio.
    double b = 2.0 * (1.0 + x);
    double d = 1.0 / b;
    double h = d, delh = d;
    const double a1 = 0.25 - mu * mu;
    double a = -a1;
    int i;
    for (i = 2; i <= kMaxIt; ++i) {
      a -= 2 * (i - 1);
      b += 2.0;
      d = 1.0 / (b + a * d);
      delh = (b * d - 1.0) * delh;
      h += delh;
      if (fabs(delh / h) < kEps) break;
    }
    if (i > kMaxIt) *ok = 0;
    h = a1 * h;
    ratio = (mu + x + 0.5 - h) * xi;
  }
  // upward ratio recurrence: K_{m+1}/K_m = 2 m / x + K_{m-1}/K_m
  double r = ratio;  // K_{mu+1}/K_mu
  for (int k = 1; k <= nl; ++k) r = 2.0 * (mu + k) * xi + 1.0 / r;
  return nu * xi - r;
}

}  // namespace bessel
}  // namespace ticc
