"""Oracle Riccati-Bessel and modified-Bessel log-derivative values.

Restates ``src/pyticc/match/bessel.py`` of the reference (scipy.special
jv/yv/jvp/yvp/kve, pinned in the reference lock file at scipy 1.17/1.18; this
container has scipy 1.15.3 whose published algorithms -- AMOS/Cephes -- are the
same routines).

Pins (tests/test_oracle_matching.py): order-0 elementary forms
(test_bessel.py:7-13), unit Wronskian at non-integer order (16-19), K
log-derivative vs kvp/kv (22-29) and at large argument (test_smatrix.py:140-145).
"""

from __future__ import annotations

import numpy as np
from scipy import special


def riccati_bessel_jy(ell: float, x: float) -> tuple[float, float, float, float]:
    """(j, n, j', n') with j = sqrt(pi x/2) J_{l+1/2}(x) (bessel.py:7-45)."""
    if x <= 0.0 or ell < 0.0:
        raise ValueError(f"x must be positive and ell must be non-negative, but got x={x}, ell={ell}")
    nu = ell + 0.5
    scale = np.sqrt(np.pi * x / 2.0)
    J = special.jv(nu, x)
    Yv = special.yv(nu, x)
    Jp = special.jvp(nu, x)
    Yp = special.yvp(nu, x)
    out = (scale * J, scale * Yv, scale * (Jp + 0.5 * J / x), scale * (Yp + 0.5 * Yv / x))
    if not np.all(np.isfinite(out)):
        raise ValueError(f"Riccati-Bessel functions and derivatives must be finite, but got {out}")
    return out


def modified_bessel_K_logD(nu: float, x: float) -> float:
    """K'_nu(x)/K_nu(x) = -K_{nu-1}/K_nu - nu/x via kve (bessel.py:90-122)."""
    if x <= 0.0 or nu < 0.0:
        raise ValueError(f"x must be positive and nu must be non-negative, but got x={x}, nu={nu}")
    v = float(-special.kve(nu - 1.0, x) / special.kve(nu, x) - nu / x)
    if not np.isfinite(v):
        raise ValueError("Modified Bessel K logarithmic derivative must be finite")
    return v


def modified_bessel_I_logD(nu: float, x: float) -> float:
    """I'_nu(x)/I_nu(x) = I_{nu-1}/I_nu - nu/x via ive (bessel.py:52-83)."""
    return float(special.ive(nu - 1.0, x) / special.ive(nu, x) - nu / x)
