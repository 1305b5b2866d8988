"""Angular quadrature and coupling coefficients (src/pyticc/basis/angle.py of the reference).

gauss_legendre_dvr restates angle.py:13-52, norm_YjK 69-97, lambda_plus
301-315.  clebsch_gordan (angle.py:217-233 uses sympy) is evaluated here from
Racah's closed formula in exact rational arithmetic, then rounded once.
"""

from __future__ import annotations

from fractions import Fraction
from functools import lru_cache
from math import factorial, isqrt

import numpy as np
from scipy.special import gammaln, lpmv, roots_legendre

from pyticc_amd._log import fail


def gauss_legendre_dvr(lower: float, upper: float, n_points: int, symmetry: bool = False):
    """GL nodes/weights on [lower, upper]; symmetry keeps half of a 2n rule with doubled weights."""
    if n_points < 1:
        raise fail(ValueError, f"n_points must be positive, but got {n_points}")
    if not np.isfinite(lower) or not np.isfinite(upper) or lower >= upper:
        raise fail(ValueError, f"Quadrature bounds must be finite and increasing, but got lower={lower}, upper={upper}")
    n_full = 2 * n_points if symmetry else n_points
    x, w = roots_legendre(n_full)
    center = 0.5 * (upper + lower)
    half = 0.5 * (upper - lower)
    full_w = half * w
    if symmetry:
        return (center + half * x[:n_points]).copy(), (2.0 * full_w[:n_points]).copy()
    return (center + half * x).copy(), full_w.copy()


def norm_YjK(j: int, K: int, x):
    """sqrt((2j+1)/2 (j-m)!/(j+m)!) P_j^m(x), m = |K|, sign -1 for odd negative K."""
    m = abs(K)
    if m > j:
        raise fail(ValueError, f"Invalid input: |K|={m} exceeds j={j}")
    factor = np.exp(0.5 * (np.log((2.0 * j + 1.0) / 2.0) + gammaln(j - m + 1.0) - gammaln(j + m + 1.0)))
    if K < 0 and m % 2 == 1:
        factor *= -1.0
    return factor * lpmv(m, j, x)


@lru_cache(maxsize=None)
def clebsch_gordan(j1: int, m1: int, j2: int, m2: int, j_couple: int) -> float:
    """<j1 m1, j2 m2 | J, m1+m2> (Condon-Shortley phase) by Racah's formula."""
    J = j_couple
    M = m1 + m2
    if abs(m1) > j1 or abs(m2) > j2 or abs(M) > J or J < abs(j1 - j2) or J > j1 + j2:
        return 0.0
    f = factorial
    pref = Fraction(
        (2 * J + 1) * f(J + j1 - j2) * f(J - j1 + j2) * f(j1 + j2 - J) * f(J + M) * f(J - M) * f(j1 - m1) * f(j1 + m1) * f(j2 - m2) * f(j2 + m2),
        f(j1 + j2 + J + 1),
    )
    total = Fraction(0)
    kmin = max(0, j2 - J - m1, j1 - J + m2)
    kmax = min(j1 + j2 - J, j1 - m1, j2 + m2)
    for k in range(kmin, kmax + 1):
        den = f(k) * f(j1 + j2 - J - k) * f(j1 - m1 - k) * f(j2 + m2 - k) * f(J - j2 + m1 + k) * f(J - j1 - m2 + k)
        total += Fraction((-1) ** k, den)
    if total == 0:
        return 0.0
    value = _sqrt_fraction(pref * total * total)
    return value if total > 0 else -value


def _sqrt_fraction(q: Fraction) -> float:
    """sqrt of a positive rational to within one ulp (integer square root at 2^-120 scale)."""
    shift = 240
    scaled = (q.numerator << shift) // q.denominator
    r = isqrt(scaled)
    return r / float(1 << (shift // 2))


def lambda_plus(j, K) -> float:
    """sqrt(j(j+1) - K(K+1))."""
    return np.sqrt(j * (j + 1) - K * (K + 1))
