"""Sine DVR and potential-optimised DVR of a diatomic monomer (src/pyticc/basis/dvr.py,
src/pyticc/basis/podvr.py of the reference): the setup step that turns a monomer
potential into the rovibrational basis (grids, E_vj, WF_vj) the channel basis and
the weighted PES rows are built from.

Sine DVR on (a, b) with n interior points x_l = a + l (b - a)/(n + 1): the
functions phi_m(x) = sqrt(2/(b - a)) sin(m pi (x - a)/(b - a)), m = 1..n, are the
FBR; the DVR <-> FBR transform is the orthogonal sine matrix
B_ml = sqrt(2/(n + 1)) sin(m l pi/(n + 1)), and the kinetic energy is diagonal in
the FBR, k_m^2/(2 mass) with k_m = m pi/(b - a), so T = B^T diag(k_m^2/(2 mass)) B
(the reference writes T in its closed form, dvr.py:66-108; both give the same
matrix, which the tests check).

PODVR: the lowest n_podvr eigenvectors of the reference vibrational Hamiltonian
span a contracted basis; the eigenvectors of the position operator in it are the
PODVR points and transform (podvr.py:13-45); each j adds j(j+1)/(2 mass r^2)
on the PODVR grid (podvr.py:85-135).

Eigenvector signs follow the reference's convention (dvr.py:112-143): each
column's first element of magnitude above 1e-4 is made positive, since the sign
of every WF_vj enters the S-matrix phases.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.linalg import eigh

from pyticc_amd._log import fail
from pyticc_amd.basis.rovib import RovibBasis

PHASE_THRESHOLD = 1.0e-4


def sine_dvr_grids(a: float, b: float, n: int) -> tuple[np.ndarray, float]:
    """Interior points x_l = a + l (b - a)/(n + 1), l = 1..n, and the uniform weight."""
    if n < 1:
        raise fail(ValueError, f"a sine DVR needs at least one point, got n = {n}")
    if not b > a:
        raise fail(ValueError, f"sine DVR interval must satisfy a < b, got ({a}, {b})")
    step = (b - a) / (n + 1)
    return a + step * np.arange(1, n + 1, dtype=np.float64), step


def sine_dvr_to_fbr(n: int) -> np.ndarray:
    """B_ml = sqrt(2/(n+1)) sin(m l pi/(n+1)): orthogonal and symmetric."""
    m = np.arange(1, n + 1, dtype=np.float64)
    return np.sqrt(2.0 / (n + 1)) * np.sin(np.pi * np.outer(m, m) / (n + 1))


def sine_dvr_kinetic(a: float, b: float, n: int, mass: float) -> np.ndarray:
    """Kinetic energy on the DVR points: T = B^T diag((m pi/(b - a))^2/(2 mass)) B."""
    if not np.isfinite(mass) or mass <= 0.0:
        raise fail(ValueError, f"mass must be positive and finite, but got {mass}")
    B = sine_dvr_to_fbr(n)
    k = np.pi * np.arange(1, n + 1, dtype=np.float64) / (b - a)
    T = (B.T * (k * k / (2.0 * mass))) @ B
    return 0.5 * (T + T.T)


def phase_fix(A: np.ndarray) -> np.ndarray:
    """Each vector (or column) signed so that its first element above 1e-4 in magnitude is positive."""
    res = np.array(A, dtype=np.float64, copy=True)
    if res.ndim == 1:
        big = np.flatnonzero(np.abs(res) > PHASE_THRESHOLD)
        if big.size == 0:
            raise fail(ValueError, "All elements are below the threshold.")
        return -res if res[big[0]] < 0.0 else res
    if res.ndim == 2:
        for c in range(res.shape[1]):
            res[:, c] = phase_fix(res[:, c])
        return res
    raise fail(ValueError, f"phase_fix takes a vector or a matrix, got shape {res.shape}")


@dataclass(frozen=True)
class SineDVR:
    """Sine-DVR basis with the eigenpairs of its vibrational Hamiltonian (columns)."""

    n_dvr: int
    interval: tuple
    grids: np.ndarray
    weights: float
    dvr_to_fbr: np.ndarray
    eigen_val: np.ndarray
    eigen_vec: np.ndarray


def build_SineDVR(a: float, b: float, n_dvr: int, mass: float, pot_func) -> SineDVR:
    """Sine DVR of H = T + V(r) on (a, b) with n_dvr points (dvr.py:218-306)."""
    x, w = sine_dvr_grids(a, b, n_dvr)
    T = sine_dvr_kinetic(a, b, n_dvr, mass)
    V = np.asarray(pot_func(x), dtype=np.float64).reshape(-1)
    if V.shape != (n_dvr,) or not np.all(np.isfinite(V)):
        raise fail(ValueError, f"monomer potential must return {n_dvr} finite values on the DVR grid")
    H = T.copy()
    H[np.diag_indices(n_dvr)] += V
    val, vec = eigh(H)
    return SineDVR(n_dvr, (float(a), float(b)), x, w, sine_dvr_to_fbr(n_dvr), val, phase_fix(vec))


def _check_qn(vmax: int, jmax: int, mass: float) -> None:
    if vmax < 0 or jmax < 0:
        raise fail(ValueError, f"vmax and jmax must be non-negative, but got vmax={vmax}, jmax={jmax}")
    if not np.isfinite(mass) or mass <= 0.0:
        raise fail(ValueError, f"mass must be positive and finite, but got {mass}")


def _rovib_on(grids: np.ndarray, H_ref: np.ndarray, vmax: int, jmax: int, mass: float) -> RovibBasis:
    """Lowest vmax + 1 states of H_ref + j(j+1)/(2 mass r^2) for every j <= jmax on a grid where r is diagonal."""
    n = grids.size
    n_v = min(vmax + 1, n)
    E = np.zeros((vmax + 1, jmax + 1))
    WF = np.zeros((n, vmax + 1, jmax + 1))
    for j in range(jmax + 1):
        H = H_ref.copy()
        H[np.diag_indices(n)] += j * (j + 1) / (2.0 * mass * grids**2)
        val, vec = eigh(H, subset_by_index=(0, n_v - 1))
        E[:n_v, j] = val
        WF[:, :n_v, j] = phase_fix(vec)
    return RovibBasis(grids=grids, E_vj=E, WF_vj=WF)


def build_RovibDVR(dvr: SineDVR, vmax: int, jmax: int, mass: float) -> RovibBasis:
    """Rovibrational states on the primitive DVR grid (dvr.py:160-215)."""
    _check_qn(vmax, jmax, mass)
    H_ref = (dvr.eigen_vec * dvr.eigen_val) @ dvr.eigen_vec.T
    return _rovib_on(dvr.grids, 0.5 * (H_ref + H_ref.T), vmax, jmax, mass)


def podvr_grids(dvr_grids: np.ndarray, dvr_wf: np.ndarray, n_podvr: int):
    """PODVR points, the DVR -> contracted transform and the PODVR -> contracted transform (podvr.py:13-45)."""
    n = min(int(n_podvr), dvr_wf.shape[1])
    if n < 1:
        raise fail(ValueError, f"n_podvr must be positive, got {n_podvr}")
    C = dvr_wf[:, :n].copy()
    X = (C.T * dvr_grids) @ C
    pts, U = eigh(0.5 * (X + X.T))
    return pts, C, phase_fix(U)


@dataclass(frozen=True)
class VibPODVR:
    """Contracted vibrational basis on the PODVR points: grids (n,), energies (n_v,), wavefunctions (n, n_v)."""

    grids: np.ndarray
    energies: np.ndarray
    wavefunctions: np.ndarray


def _H_podvr(U: np.ndarray, eig: np.ndarray) -> np.ndarray:
    """The reference Hamiltonian, diagonal in the contracted basis (eigenvalues eig),
    on the PODVR points: H[a, b] = sum_i U[i, a] eig_i U[i, b]."""
    H = (U.T * eig[: U.shape[0]]) @ U
    return 0.5 * (H + H.T)


def build_VibPODVR(dvr: SineDVR, n_podvr: int, vmax: int) -> VibPODVR:
    """The lowest vmax + 1 vibrational states on the PODVR points (podvr.py:48-82, 170-186)."""
    pts, _, U = podvr_grids(dvr.grids, dvr.eigen_vec, n_podvr)
    n_keep = min(vmax + 1, pts.size)
    val, vec = eigh(_H_podvr(U, dvr.eigen_val))
    E = np.zeros(vmax + 1)
    WF = np.zeros((pts.size, vmax + 1))
    E[:n_keep] = val[:n_keep]
    WF[:, :n_keep] = phase_fix(vec)[:, :n_keep]
    return VibPODVR(grids=pts, energies=E, wavefunctions=WF)


def build_RovibPODVR(dvr: SineDVR, n_podvr: int, vmax: int, jmax: int, mass: float) -> RovibBasis:
    """Rovibrational states on the PODVR points (podvr.py:85-135, 189-207)."""
    _check_qn(vmax, jmax, mass)
    pts, _, U = podvr_grids(dvr.grids, dvr.eigen_vec, n_podvr)
    return _rovib_on(pts, _H_podvr(U, dvr.eigen_val), vmax, jmax, mass)
