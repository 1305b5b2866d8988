"""Oracle asymptotic matching: BF->SF transform, closed-channel elimination, K->S.

Restates ``src/pyticc/match/asymptotic.py:16-110`` and
``src/pyticc/match/smatrix.py:10-223`` of the reference, plus two quantities the
reference does not define (unitarity defect, partial cross sections) written
from their textbook closed forms.

Pins (tests/test_oracle_matching.py): free channel S=1 (test_smatrix.py:42-53),
free propagation + matching S=1 through the oracle LDMD (56-77), K->S closed form
and unitarity (80-95), closed channel eliminated before the open block (98-112),
elimination == full reaction matrix (115-137), capture S=2/3 (148-161);
B^T U B = diag(L(L+1)), L=[1,3], last row >= 0 (test_asymptotic.py:28-38),
non-integer L = sqrt(40.25)-1/2 (60-70).  Cross sections: a one-channel
hard-sphere-like case against sin^2 of the phase shift, and detailed balance
k_i^2 g_i sigma_if = k_f^2 g_f sigma_fi on a random unitary symmetric S.
"""

from __future__ import annotations

import numpy as np

from oracle.bessel import modified_bessel_K_logD, riccati_bessel_jy


def bmat_bf_to_sf(U: np.ndarray, ladder: np.ndarray, j_couple: np.ndarray, Jtot: int):
    """Per-ladder eigh of U with the reference sign rule (asymptotic.py:16-75).

    ladder: (N,) integer label of (mis_X, mis_Y, j_couple); groups keep the
    first-appearance order of the channels.
    """
    n = U.shape[0]
    B = np.zeros((n, n))
    L = np.empty(n)
    order: dict[int, list[int]] = {}
    for i, g in enumerate(np.asarray(ladder).tolist()):
        order.setdefault(g, []).append(i)
    for pos in order.values():
        p = np.asarray(pos)
        w, v = np.linalg.eigh(U[np.ix_(p, p)])
        if np.min(w) < -1.0e-10:
            raise ValueError(f"Centrifugal matrix has a negative eigenvalue {np.min(w)}")
        j12 = int(j_couple[p[0]])
        phase = (-1) ** (j12 + min(Jtot, j12))
        v = v * np.where(v[-1] * phase < 0.0, -1.0, 1.0)
        B[np.ix_(p, p)] = v
        L[p] = np.sqrt(np.maximum(w, 0.0) + 0.25) - 0.5
    return B, L


def transform_bf_to_sf(Y: np.ndarray, B: np.ndarray) -> np.ndarray:
    """Y_SF = B^T Y B (asymptotic.py:82-110)."""
    return np.einsum("ia,...ij,jb->...ab", B, Y, B, optimize=True)


def reference_functions(energy, Rmatch, mu, E_int, L):
    """Open J,N,J',N' and closed Gamma (smatrix.py:10-79)."""
    diff = energy - np.asarray(E_int, dtype=np.float64)
    if np.any(diff == 0.0):
        raise ValueError(f"Asymptotic matching is undefined exactly at a channel threshold Etot={energy}")
    op = np.flatnonzero(diff > 0.0)
    cl = np.flatnonzero(diff < 0.0)
    J = np.empty(op.size)
    N = np.empty(op.size)
    Jp = np.empty(op.size)
    Np = np.empty(op.size)
    G = np.empty(cl.size)
    for a, i in enumerate(op):
        k = float(np.sqrt(2.0 * mu * abs(diff[i])))
        j, n, jd, nd = riccati_bessel_jy(float(L[i]), k * Rmatch)
        J[a] = j / np.sqrt(k)
        N[a] = n / np.sqrt(k)
        Jp[a] = np.sqrt(k) * jd
        Np[a] = np.sqrt(k) * nd
    for c, i in enumerate(cl):
        k = float(np.sqrt(2.0 * mu * abs(diff[i])))
        G[c] = 0.5 / Rmatch + k * modified_bessel_K_logD(float(L[i] + 0.5), k * Rmatch)
    return op, cl, J, N, Jp, Np, G


def _solve(A, B, inverse):
    """A^-1 B by a linear solve, or (inverse=True) through the explicit inverse --
    two FP64 orderings of the same step."""
    return np.linalg.inv(A) @ B if inverse else np.linalg.solve(A, B)


def open_reaction_matrix(Y, op, cl, J, N, Jp, Np, G, inverse: bool = False):
    """Y_eff = Y_oo - Y_oc (Y_cc - G)^-1 Y_co; K = -(Y_eff N - N')^-1 (Y_eff J - J') (smatrix.py:86-132)."""
    Yo = Y[np.ix_(op, op)]
    if cl.size:
        Yoc = Y[np.ix_(op, cl)]
        Yco = Y[np.ix_(cl, op)]
        Ycc = Y[np.ix_(cl, cl)] - np.diag(G)
        Yo = Yo - Yoc @ _solve(Ycc, Yco, inverse)
    left = Yo * N[None, :] - np.diag(Np)
    right = Yo * J[None, :] - np.diag(Jp)
    return np.asarray(-_solve(left, right, inverse), dtype=np.complex128)


def get_smat(Y, Rmatch, energies, mu, E_int, L, inverse: bool = False):
    """Ragged tuple of open-open S matrices (smatrix.py:139-223).  inverse=True
    forms every solve through an explicit inverse: another FP64 ordering, whose
    distance from the solve-based S is the matching step's own noise floor."""
    out = []
    for e, energy in enumerate(np.asarray(energies, dtype=np.float64)):
        op, cl, J, N, Jp, Np, G = reference_functions(float(energy), Rmatch, mu, E_int, L)
        if op.size == 0:
            out.append(np.empty((0, 0), dtype=np.complex128))
            continue
        K = open_reaction_matrix(Y[e], op, cl, J, N, Jp, Np, G, inverse)
        eye = np.eye(op.size, dtype=np.complex128)
        out.append(_solve(eye + 1.0j * K, eye - 1.0j * K, inverse))
    return tuple(out)


def unitarity_defect(S: np.ndarray) -> float:
    """max |S^dagger S - I| (the check of test_smatrix.py:95, test_high_level.py:231)."""
    if S.size == 0:
        return 0.0
    return float(np.max(np.abs(S.conj().T @ S - np.eye(S.shape[0]))))


def partial_cross_sections(S, open_levels, k_open, degeneracy, Jtot):
    """sigma^J_{a->b} = pi (2J+1) / (k_a^2 g_a) sum_{c in a, c' in b} |delta_cc' - S_c'c|^2.

    Textbook (Arthurs-Dalgarno) partial cross section of one (J, parity) block;
    the reference has no such function.  open_levels: (n_open,) level label per
    open channel in S order; k_open: (n_open,) wave numbers; degeneracy: dict
    level -> g.  Returns (levels, sigma[a, b]).
    """
    levels = sorted(set(np.asarray(open_levels).tolist()))
    T2 = np.abs(np.eye(S.shape[0]) - S) ** 2  # T2[c', c] = |delta - S_c'c|^2
    sig = np.zeros((len(levels), len(levels)))
    lab = np.asarray(open_levels)
    for a, la in enumerate(levels):
        ca = np.flatnonzero(lab == la)
        ka = float(k_open[ca[0]])
        for b, lb in enumerate(levels):
            cb = np.flatnonzero(lab == lb)
            sig[a, b] = np.pi * (2 * Jtot + 1) / (ka**2 * degeneracy[la]) * float(T2[np.ix_(cb, ca)].sum())
    return levels, sig
