"""Readers of the committed golden fixtures (NumPy only: bench.py's parity check
uses them without importing the oracle)."""

from __future__ import annotations

import numpy as np


def rotor_rotor_N800_reference(d, i):
    """(rows, cols, values) of the fixture's stored elements of S at its i-th energy."""
    n = int(d["n_open"][i])
    if f"S{i}_upper" in d.files:
        r, c = np.triu_indices(n)
        return r, c, d[f"S{i}_upper"]
    if f"S{i}_cols" in d.files:
        cols = d[f"S{i}_colidx"]
        r = np.concatenate([np.repeat(np.arange(n), cols.size), np.arange(n)])
        c = np.concatenate([np.tile(cols, n), np.arange(n)])
        v = np.concatenate([d[f"S{i}_cols"].ravel(), d[f"S{i}_diag"]])
        return r, c, v
    return np.zeros(0, int), np.zeros(0, int), np.zeros(0, complex)


def n800_tolerance(d, i, bar=1.0e-10):
    """Per-energy |dS| tolerance at full size for S computed from a propagated Y
    by the oracle's matching: the 1e-10 bar, or twice the FP64 noise floor of the
    contraction and sweep recorded with the fixture (the largest distance between
    the oracle and its other orderings) where that floor exceeds half of the bar."""
    return max(bar, 2.0 * float(d["floor"][i]))


def n800_match_tolerance(d, i, bar=1.0e-10):
    """Per-energy tolerance for two matchings of the same Y: the 1e-10 bar, or
    twice the matching step's own floor recorded with the fixture."""
    fm = float(d["floor_match"][i]) if "floor_match" in d.files else 0.0
    return max(bar, 2.0 * fm)


def n800_total_tolerance(d, i, bar=1.0e-10):
    """Bound on the distance of a device S (its own sweep and matching) from the
    fixture: the two parts' tolerances added."""
    return n800_tolerance(d, i, bar) + n800_match_tolerance(d, i, bar)
