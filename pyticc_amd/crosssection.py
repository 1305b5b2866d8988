"""Partial state-to-state cross sections of one (J, parity) block.

The reference has no cross-section function (its matching ends at the S matrix,
src/pyticc/match/smatrix.py:139-223, and result.py:39-86 stores S only).  This
is the textbook partial integral cross section of a rigid-rotor collision
(Arthurs and Dalgarno, Proc. R. Soc. A 256, 540 (1960); for two rotors Green,
J. Chem. Phys. 62, 2271 (1975)):

    sigma^{Jp}(a -> b) = pi (2J + 1) / (k_a^2 g_a) * sum_{c in a, c' in b} |delta_{c'c} - S_{c'c}|^2,

where a, b are the monomer levels (the (X, Y) internal states of the channels),
c runs over the open channels of level a, k_a = sqrt(2 mu (E - E_a)) and
g_a = (2 j_X + 1)(2 j_Y + 1) is the level's m-degeneracy (j_Y = 0 for an atom).
Summing over J and parity gives the integral cross section.  S[c', c] is the
amplitude from incoming channel c to outgoing channel c' in the open-channel
order flatnonzero(E_int < E) that ScatteringResult.Smat uses; the BF -> SF
transform only mixes channels of one (j_X, j_Y, j12) ladder, so each position
keeps its channel's level.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from pyticc_amd._log import fail


@dataclass(frozen=True)
class PartialCrossSections:
    """sigma[a, b] = sigma^{Jp}(levels[a] -> levels[b]) in bohr^2 at one energy."""

    Etot: float
    levels: tuple   # ((mis_X, mis_Y), ...) open levels, ascending threshold energy
    E_level: np.ndarray
    sigma: np.ndarray


def level_of(channel):
    return (channel.mis_X, channel.mis_Y)


def degeneracy(level) -> int:
    mis_X, mis_Y = level
    return (2 * int(mis_X.j) + 1) * (2 * int(mis_Y.j) + 1)


def partial_cross_sections(result, reduced_mass: float | None = None) -> tuple:
    """One PartialCrossSections per energy of an exact-CC ScatteringResult."""
    mu = result.reduced_mass if reduced_mass is None else reduced_mass
    if mu is None or not mu > 0.0:
        raise fail(ValueError, "partial_cross_sections needs a positive reduced_mass")
    basis = result.basis
    Jtot = int(basis.Jtot)
    E_int = basis.E_int
    chan_level = [level_of(c) for c in basis]
    out = []
    for e, (E, S) in enumerate(zip(np.asarray(result.Etot, dtype=np.float64), result.Smat)):
        open_idx = np.flatnonzero(E_int < E)
        if S.shape != (open_idx.size, open_idx.size):
            raise fail(ValueError, f"S at energy index {e} has shape {S.shape}, expected {(open_idx.size,) * 2}")
        labels = [chan_level[i] for i in open_idx]
        levels = []
        for lab in labels:
            if lab not in levels:
                levels.append(lab)
        E_lev = np.array([E_int[open_idx[labels.index(lv)]] for lv in levels])
        order = np.argsort(E_lev, kind="stable")
        levels = [levels[i] for i in order]
        E_lev = E_lev[order]
        pos = {lv: i for i, lv in enumerate(levels)}
        onehot = np.zeros((open_idx.size, len(levels)))
        onehot[np.arange(open_idx.size), [pos[lab] for lab in labels]] = 1.0
        T2 = np.abs(np.eye(open_idx.size) - S) ** 2  # T2[c', c]
        summed = onehot.T @ T2 @ onehot  # summed[b, a] = sum_{c in a, c' in b} T2[c', c]
        k2 = 2.0 * mu * (E - E_lev)
        g = np.array([degeneracy(lv) for lv in levels], dtype=np.float64)
        sigma = np.pi * (2 * Jtot + 1) / (k2 * g)[:, None] * summed.T
        out.append(PartialCrossSections(Etot=float(E), levels=tuple(levels), E_level=E_lev, sigma=sigma))
    return tuple(out)
