"""Oracle restatement of the Manolopoulos LDMD propagator.

Restates ``src/pyticc/propagation/logd.py`` of the reference.

Pins (tests/test_oracle_ldmd.py):
  * constant single-channel W: exact tanh/tan propagator, rtol 1e-13
    (reference KAT tests/unit/test_logd.py:46-55, 83-95);
  * capture incoming wave is invariant (test_logd.py:58-67);
  * constant *coupled* W: converges to the exact eigenbasis propagator with
    O(h^4) error (a dropped Q term or wrong Q_mid sign breaks the order);
  * batched sweep == sequential single-sector calls (test_logd.py:98-144);
  * initial conditions (test_logd.py:34-43);
  * the other FP64 orderings of the sector step (explicit inverses, blocked
    Gauss-Jordan) agree with the solve-based step to rounding, and the
    Gauss-Jordan inverse equals LAPACK's on random, permutation and zero-diagonal
    matrices at every block size.
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


def gauss_jordan_inverse(A, nb: int = 32):
    """Inverse by blocked Gauss-Jordan elimination with partial pivoting: columns
    in blocks of nb; in each block the pivot of every column is the largest |a|
    among the rows at positions >= the column (the first on ties, as LAPACK's
    idamax), found by in-place Gauss-Jordan on the block's column panel, and the
    block's transformation is then applied to the other columns at once
    (pivot rows B^-1 R, other rows A_i - A_i[panel] B^-1 R).  Row interchanges
    make the result the inverse of the row-permuted matrix; its columns are
    gathered back at the end.  Another FP64 ordering of the inversions in
    logd.py:168-194 (the device inverts this way, with its own blocking)."""
    X = np.array(A, dtype=np.float64, copy=True)
    n = X.shape[0]
    perm = np.arange(n)
    for k0 in range(0, n, nb):
        c = slice(k0, min(n, k0 + nb))
        w = c.stop - k0
        P = X[:, c].copy()
        rows = np.arange(n)  # panel row order: position -> row of X
        for j in range(w):
            k = k0 + j
            p = k + int(np.argmax(np.abs(P[k:, j])))
            if p != k:
                P[[k, p]] = P[[p, k]]
                rows[[k, p]] = rows[[p, k]]
            piv = P[k, j]
            P[k] /= piv
            P[k, j] = 1.0 / piv
            f = P[:, j].copy()
            f[k] = 0.0
            P -= np.outer(f, P[k])
            P[:, j] = -f / piv
            P[k, j] = 1.0 / piv
        moved = np.nonzero(rows != np.arange(n))[0]
        X[moved] = X[rows[moved]]
        perm = perm[rows]
        R = X[k0:c.stop].copy()
        # (the products also cover the panel columns, which P then overwrites)
        X[:k0] += P[:k0] @ R
        X[c.stop:] += P[c.stop:] @ R
        X[k0:c.stop] = P[k0:c.stop] @ R
        X[:, c] = P
    out = np.empty_like(X)
    out[:, perm] = X
    return out


def sector_step_gj(Y, h, ref, Qs, Qm, Qe):
    """The sector with both inverses formed by blocked Gauss-Jordan
    (gauss_jordan_inverse) and then scaled: a third FP64 ordering of
    logd.py:168-194, of the same algorithm class as the device's, whose distance
    from the solve-based step bounds what an end-to-end comparison can resolve."""
    y1, y2 = reference_values(h, ref)
    n = Y.shape[-1]
    idx = np.arange(n)
    A1 = Y + Qs
    A1[idx, idx] += y1
    X1 = gauss_jordan_inverse(A1) * y2[None, :]
    Ymid = Qm.copy()
    Ymid[idx, idx] += y1
    Ymid -= y2[:, None] * X1
    A2 = Ymid + Qm
    A2[idx, idx] += y1
    X2 = gauss_jordan_inverse(A2) * y2[None, :]
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
