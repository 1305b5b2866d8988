"""BF centrifugal / Coriolis matrix (src/pyticc/matrix/centrifugal.py:27-85)."""

from __future__ import annotations

import numpy as np

from pyticc_amd.basis.angle import lambda_plus


def get_Umat_BF(basis, channel_indices=None, *, coriolis: bool = True) -> np.ndarray:
    """U_ii = J(J+1) + j12(j12+1) - 2K^2; U_{K,K+1} = -(sqrt2 if K=0) lambda+(J,K) lambda+(j12,K) within a ladder."""
    indices = tuple(range(basis.n_channel)) if channel_indices is None else tuple(channel_indices)
    channels = tuple(basis[i] for i in indices)
    U = np.zeros((len(channels), len(channels)))
    groups: dict = {}
    for li, ch in enumerate(channels):
        groups.setdefault((ch.mis_X, ch.mis_Y, ch.j_couple), {})[ch.K] = li
        U[li, li] = basis.Jtot * (basis.Jtot + 1) + ch.j_couple * (ch.j_couple + 1) - 2 * ch.K**2
    if not coriolis:
        return U
    for K_to_index in groups.values():
        for K, li in K_to_index.items():
            nxt = K_to_index.get(K + 1)
            if nxt is None:
                continue
            ch = channels[li]
            c = -(np.sqrt(2.0) if K == 0 else 1.0) * lambda_plus(basis.Jtot, K) * lambda_plus(ch.j_couple, K)
            U[li, nxt] = c
            U[nxt, li] = c
    return U


def ladder_labels(basis, channel_indices=None) -> np.ndarray:
    """Integer label of each channel's (mis_X, mis_Y, j12) ladder, in first-appearance order."""
    indices = tuple(range(basis.n_channel)) if channel_indices is None else tuple(channel_indices)
    seen: dict = {}
    out = np.empty(len(indices), dtype=np.int64)
    for li, i in enumerate(indices):
        ch = basis[i]
        out[li] = seen.setdefault((ch.mis_X, ch.mis_Y, ch.j_couple), len(seen))
    return out
