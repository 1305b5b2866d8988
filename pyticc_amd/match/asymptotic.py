"""BF -> SF asymptotic transform (src/pyticc/match/asymptotic.py:16-110).

The per-ladder eigendecomposition of U is one-time setup on the host (as in
the reference); the transform of the propagated Y is fused into the device
matching kernel (ticc_get_smat).
"""

from __future__ import annotations

import numpy as np

from pyticc_amd._log import fail
from pyticc_amd.matrix.centrifugal import get_Umat_BF


def get_Bmat_BF_to_SF(basis, channel_indices=None):
    """(B, L) with B^T U B = diag(L(L+1)); eigenvector sign: last row x (-1)^(j12+min(J,j12)) >= 0."""
    indices = tuple(range(basis.n_channel)) if channel_indices is None else tuple(channel_indices)
    if not indices:
        raise fail(ValueError, "At least one channel is required for the BF-to-SF transformation")
    channels = tuple(basis[i] for i in indices)
    U = get_Umat_BF(basis, indices)
    n = len(indices)
    B = np.zeros((n, n))
    L = np.empty(n)
    groups: dict = {}
    for li, ch in enumerate(channels):
        groups.setdefault((ch.mis_X, ch.mis_Y, ch.j_couple), []).append(li)
    for pos_list in groups.values():
        pos = np.asarray(pos_list, dtype=np.int64)
        w, v = np.linalg.eigh(U[np.ix_(pos, pos)])
        if np.min(w) < -1.0e-10:
            raise fail(ValueError, f"Centrifugal matrix has a negative eigenvalue {np.min(w)}")
        ch = channels[pos_list[0]]
        phase = (-1) ** (ch.j_couple + min(basis.Jtot, ch.j_couple))
        v = v * np.where(v[-1] * phase < 0.0, -1.0, 1.0)
        B[np.ix_(pos, pos)] = v
        L[pos] = np.sqrt(np.maximum(w, 0.0) + 0.25) - 0.5
    return B, L


def transform_logD_BF_to_SF(Ymat, Bmat):
    """Y_SF = B^T Y B for host arrays (result packaging; the device path fuses it into matching)."""
    Y = np.asarray(Ymat)
    B = np.asarray(Bmat, dtype=np.float64)
    if B.ndim != 2 or B.shape[0] != B.shape[1]:
        raise fail(ValueError, f"Bmat must be square, but got shape={B.shape}")
    if Y.ndim < 2 or Y.shape[-2:] != B.shape:
        raise fail(ValueError, f"Ymat must end with matrix shape {B.shape}, but got shape={Y.shape}")
    return np.einsum("ia,...ij,jb->...ab", B, Y, B, optimize=True)
