"""Oracle restatement of the Manolopoulos LDMD propagator.

Restates ``src/pyticc/propagation/logd.py`` of the reference.

Pins (tests/test_oracle_ldmd.py):
  * constant single-channel W: exact tanh/tan propagator, rtol 1e-13
    (reference KAT tests/unit/test_logd.py:46-55, 83-95);
  * capture incoming wave is invariant (test_logd.py:58-67);
  * constant *coupled* W: converges to the exact eigenbasis propagator with
    O(h^4) error (a dropped Q term or wrong Q_mid sign breaks the order);
  * batched sweep == sequential single-sector calls (test_logd.py:98-144);
  * initial conditions (test_logd.py:34-43).
"""

from __future__ import annotations

import numpy as np

SERIES_THRESHOLD = 1.0e-5  # logd.py:120


def initialize_inelastic(W: np.ndarray) -> np.ndarray:
    """Y_ij(R0) = delta_ij sqrt|W_ii|  (logd.py:48-65)."""
    d = np.real(np.diagonal(np.asarray(W), axis1=-2, axis2=-1))
    v = np.sqrt(np.abs(d))
    return v[..., :, None] * np.eye(v.shape[-1])


def initialize_capture(W: np.ndarray) -> np.ndarray:
    """Y_ii = sqrt(W_ii) if W_ii >= 0 else -i sqrt(-W_ii)  (logd.py:72-91)."""
    d = np.real(np.diagonal(np.asarray(W), axis1=-2, axis2=-1))
    v = np.where(d >= 0.0, np.sqrt(np.maximum(d, 0.0)), -1.0j * np.sqrt(np.maximum(-d, 0.0)))
    return v.astype(np.complex128)[..., :, None] * np.eye(v.shape[-1])


def reference_values(h: float, p2: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Diagonal reference propagators y1, y2 from p^2 = diag(W_mid) (logd.py:98-132)."""
    p2 = np.asarray(p2, dtype=np.float64)
    mag = np.sqrt(np.abs(p2))
    arg = mag * h
    small = np.abs(arg) < SERIES_THRESHOLD
    safe = np.where(small, 1.0, arg)
    with np.errstate(divide="ignore", invalid="ignore"):
        y1p = mag / np.tanh(safe)
        y1n = mag / np.tan(safe)
        y2p = mag / np.sinh(safe)
        y2n = mag / np.sin(safe)
    y1s = 1.0 / h + p2 * h / 3.0 - p2**2 * h**3 / 45.0
    y2s = 1.0 / h - p2 * h / 6.0 + 7.0 * p2**2 * h**3 / 360.0
    y1 = np.where(small, y1s, np.where(p2 >= 0.0, y1p, y1n))
    y2 = np.where(small, y2s, np.where(p2 >= 0.0, y2p, y2n))
    return y1, y2


def correction_matrices(h: float, Ws: np.ndarray, Wm: np.ndarray, We: np.ndarray):
    """Energy-independent sector corrections (logd.py:143-164).

    Returns (ref, Q_start, Q_mid, Q_end) with ref = diag(W_mid),
    Q_mid = (4/h)[(I - h^2 T_mid / 6)^-1 - I]  (the code's sign, logd.py:162).
    """
    n = Wm.shape[-1]
    eye = np.eye(n)
    ref = np.real(np.diag(Wm)).copy()
    R = np.diag(ref)
    Qs = h * (Ws - R) / 3.0
    Qm = 4.0 / h * (np.linalg.solve(eye - h**2 * (Wm - R) / 6.0, eye) - eye)
    Qe = h * (We - R) / 3.0
    return ref, Qs, Qm, Qe


def sector_step(Y, h, ref, Qs, Qm, Qe):
    """One two-half-step LDMD sector for one energy (logd.py:168-194)."""
    y1, y2 = reference_values(h, ref)
    n = Y.shape[-1]
    idx = np.arange(n)
    A1 = Y + Qs
    A1 = A1.astype(np.result_type(A1, y1))
    A1[idx, idx] += y1
    X1 = np.linalg.solve(A1, np.diag(y2).astype(A1.dtype))
    Ymid = Qm.astype(A1.dtype).copy()
    Ymid[idx, idx] += y1
    Ymid -= y2[:, None] * X1
    A2 = Ymid + Qm
    A2[idx, idx] += y1
    X2 = np.linalg.solve(A2, np.diag(y2).astype(A2.dtype))
    Yend = Qe.astype(A2.dtype).copy()
    Yend[idx, idx] += y1
    Yend -= y2[:, None] * X2
    return 0.5 * (Yend + Yend.T)


def sector_step_inverse(Y, h, ref, Qs, Qm, Qe):
    """The same sector with A^-1 formed explicitly (LAPACK getri) and then scaled,
    instead of solved against diag(y2) -- an equally valid FP64 ordering of
    logd.py:168-194.  The distance between the S-matrices of the two orderings is
    the FP64 noise floor of an end-to-end comparison."""
    y1, y2 = reference_values(h, ref)
    n = Y.shape[-1]
    idx = np.arange(n)
    A1 = Y + Qs
    A1[idx, idx] += y1
    X1 = np.linalg.inv(A1) * y2[None, :]
    Ymid = Qm.copy()
    Ymid[idx, idx] += y1
    Ymid -= y2[:, None] * X1
    A2 = Ymid + Qm
    A2[idx, idx] += y1
    X2 = np.linalg.inv(A2) * y2[None, :]
    Yend = Qe.copy()
    Yend[idx, idx] += y1
    Yend -= y2[:, None] * X2
    return 0.5 * (Yend + Yend.T)


def propagate(Y0, energies, mu, half_steps, Wb_start, Wb_mid, Wb_end):
    """Energy batch x sector scan (logd.py:288-330, 340-398).

    Y0: (n_E, N, N); energies (n_E,); half_steps (n_S,); W_base_*: (n_S, N, N).
    W(R;E) = W_base(R) - 2 mu E I enters only through the reference values.
    """
    Y = np.array(Y0, copy=True)
    shift = 2.0 * mu * np.asarray(energies, dtype=np.float64)
    for s, h in enumerate(np.asarray(half_steps, dtype=np.float64)):
        ref, Qs, Qm, Qe = correction_matrices(h, Wb_start[s], Wb_mid[s], Wb_end[s])
        for e in range(Y.shape[0]):
            Y[e] = sector_step(Y[e], h, ref - shift[e], Qs, Qm, Qe)
    return Y


def constant_logd(Y0: complex, W: float, length: float) -> complex:
    """Exact single-channel log-derivative propagator for constant W."""
    if W > 0.0:
        p = np.sqrt(W)
        t = np.tanh(p * length)
        return p * (Y0 + p * t) / (p + Y0 * t)
    k = np.sqrt(-W)
    t = np.tan(k * length)
    return k * (Y0 - k * t) / (k + Y0 * t)
