"""Result containers (src/pyticc/result.py:27-86)."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True, slots=True)
class Timing:
    wall_seconds: float
    cpu_seconds: float

    def __str__(self) -> str:
        return f"wall={self.wall_seconds:.3f} s, CPU={self.cpu_seconds:.3f} s"


@dataclass(frozen=True)
class ScatteringResult:
    """Exact result of one (J, parity) block; S[i] has shape (n_open[i], n_open[i])."""

    basis: object
    Etot: np.ndarray
    Y_propagated: np.ndarray
    asymptotic_transform: np.ndarray
    L: np.ndarray
    Smat: tuple
    timing: Timing | None = None

    @property
    def molecule_exchange(self) -> int:
        return self.basis.molecule_exchange

    @property
    def open_closed(self):
        return self.basis.open_closed(self.Etot)

    @property
    def Y_asymptotic(self) -> np.ndarray:
        return np.einsum("ia,...ij,jb->...ab", self.asymptotic_transform, self.Y_propagated, self.asymptotic_transform, optimize=True)

    @property
    def open_channel_indices(self) -> tuple:
        return tuple(np.asarray(np.flatnonzero(m), dtype=np.int64) for m in self.open_closed.open_mask)
