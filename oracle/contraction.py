"""Oracle W(R) contraction and radial-matrix assembly.

Restates ``src/pyticc/matrix/interaction/common.py:49-62, 97-146`` (per-K
contraction, symmetrised, scattered block-diagonally) and
``src/pyticc/matrix/radial.py:7-57`` / ``src/pyticc/propagation/runner.py:265-271``
(W_base = U/R^2 + 2 mu V + 2 mu diag(E_int)).

Pins (tests/test_oracle_contraction.py): constant V -> V*I for the atom-diatom
and diatom-diatom bases (test_interaction.py:116-151), block order
(154-165), radial batch (168-179), einsum GEMM identity
(test_vmat_matmul.py:16-30), W signs (test_radial.py:6-24).
"""

from __future__ import annotations

import numpy as np


def contract_block(Br: np.ndarray, Bi: np.ndarray | None, norm: float, v: np.ndarray) -> np.ndarray:
    """norm (Br diag(v) Br^T + Bi diag(v) Bi^T), then 1/2 (X + X^T) (common.py:49-62)."""
    X = (Br * v[None, :]) @ Br.T
    if Bi is not None:
        X += (Bi * v[None, :]) @ Bi.T
    X *= norm
    return 0.5 * (X + X.T)


def contract(blocks, n_channel: int, norm: float, V: np.ndarray) -> np.ndarray:
    """Block-diagonal V(R) for a radial batch (common.py:97-146).

    blocks: sequence of (indices (N_K,), B_real (N_K, n_ang), B_imag or None).
    V: (n_R, n_ang).  Cross-K elements stay exactly zero.
    """
    V = np.atleast_2d(np.asarray(V, dtype=np.float64))
    out = np.zeros((V.shape[0], n_channel, n_channel))
    for idx, Br, Bi in blocks:
        idx = np.asarray(idx)
        for r in range(V.shape[0]):
            out[r][np.ix_(idx, idx)] = contract_block(Br, Bi, norm, V[r])
    return out


def contract_split(blocks, n_channel: int, norm: float, V: np.ndarray) -> np.ndarray:
    """contract() with each angular sum taken in two halves (the first half of the
    quadrature points, then the rest) and the partial matrices added: the same
    arithmetic in another FP64 summation order, as the device's tiled contraction
    sums in its own order (used to bound what an end-to-end comparison resolves)."""
    V = np.atleast_2d(np.asarray(V, dtype=np.float64))
    out = np.zeros((V.shape[0], n_channel, n_channel))
    for idx, Br, Bi in blocks:
        idx = np.asarray(idx)
        h = Br.shape[1] // 2
        for r in range(V.shape[0]):
            v = V[r]
            X = (Br[:, :h] * v[None, :h]) @ Br[:, :h].T + (Br[:, h:] * v[None, h:]) @ Br[:, h:].T
            if Bi is not None:
                X += (Bi[:, :h] * v[None, :h]) @ Bi[:, :h].T + (Bi[:, h:] * v[None, h:]) @ Bi[:, h:].T
            X *= norm
            out[r][np.ix_(idx, idx)] = 0.5 * (X + X.T)
    return out


def w_base(R: np.ndarray, mu: float, E_int: np.ndarray, U: np.ndarray, Vmat: np.ndarray) -> np.ndarray:
    """W_base(R) = U/R^2 + 2 mu V(R) + 2 mu diag(E_int)  (runner.py:265-271)."""
    R = np.asarray(R, dtype=np.float64)
    W = U[None, :, :] / R[:, None, None] ** 2 + 2.0 * mu * Vmat
    n = U.shape[0]
    W[:, np.arange(n), np.arange(n)] += 2.0 * mu * np.asarray(E_int)[None, :]
    return W


def w_full(R: float, Etot: float, mu: float, E_int, U, Vmat) -> np.ndarray:
    """W(R;E) = U/R^2 + 2 mu V + 2 mu diag(E_int - E)  (radial.py:7-57)."""
    if R <= 0.0:
        raise ValueError(f"R must be positive, but got R={R}")
    W = U / R**2 + 2.0 * mu * Vmat
    n = U.shape[0]
    W[np.arange(n), np.arange(n)] += 2.0 * mu * (np.asarray(E_int) - Etot)
    return W
