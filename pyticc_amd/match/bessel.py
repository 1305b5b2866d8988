"""Reference functions evaluated on the device (replaces src/pyticc/match/bessel.py:7-45, 90-122)."""

from __future__ import annotations

import numpy as np
import torch

from pyticc_amd import _backend, _device
from pyticc_amd._log import fail


def riccati_bessel_batch(ell, x, device=None) -> np.ndarray:
    """(n, 4) array of j, n, j', n' (Riccati-Bessel, order ell) at x, computed by the HIP kernel."""
    dev = _device.require_gpu(0 if device is None else device)
    ell_d = _device.to_device(np.atleast_1d(ell), dev)
    x_d = _device.to_device(np.atleast_1d(x), dev)
    out = torch.empty((x_d.numel(), 4), dtype=torch.float64, device=dev)
    _backend.call("ticc_bessel_eval", x_d.numel(), _device.ptr(ell_d), _device.ptr(x_d), _device.ptr(out), None, None,
                  _device.stream_handle(dev))
    return _device.to_host(out)


def k_logd_batch(nu, x, device=None) -> np.ndarray:
    """K'_nu(x)/K_nu(x) for arrays nu, x on the device."""
    dev = _device.require_gpu(0 if device is None else device)
    nu_d = _device.to_device(np.atleast_1d(nu), dev)
    x_d = _device.to_device(np.atleast_1d(x), dev)
    out = torch.empty(x_d.numel(), dtype=torch.float64, device=dev)
    _backend.call("ticc_bessel_eval", x_d.numel(), None, _device.ptr(x_d), None, _device.ptr(nu_d), _device.ptr(out),
                  _device.stream_handle(dev))
    return _device.to_host(out)


def riccati_bessel_jy(ell: float, x: float):
    """(j, n, j', n') with j = sqrt(pi x/2) J_{ell+1/2}(x)."""
    if x <= 0.0 or ell < 0.0:
        raise fail(ValueError, f"x must be positive and ell must be non-negative, but got x={x}, ell={ell}")
    v = riccati_bessel_batch([ell], [x])[0]
    if not np.all(np.isfinite(v)):
        raise fail(ValueError, f"Riccati-Bessel functions and derivatives must be finite, but got {v}")
    return tuple(float(a) for a in v)


def modified_bessel_K_logD(nu: float, x: float) -> float:
    """K'_nu(x)/K_nu(x)."""
    if x <= 0.0 or nu < 0.0:
        raise fail(ValueError, f"x must be positive and nu must be non-negative, but got x={x}, nu={nu}")
    v = float(k_logd_batch([nu], [x])[0])
    if not np.isfinite(v):
        raise fail(ValueError, f"Modified Bessel K logarithmic derivative must be finite, but got {v}")
    return v
