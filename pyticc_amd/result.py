"""Result containers (src/pyticc/result.py:27-86, 146-238)."""

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
    reduced_mass: float | None = None  # this build's addition: cross sections need k_a

    @property
    def molecule_exchange(self) -> int:
        return self.basis.molecule_exchange

    @property
    def open_closed(self):
        return self.basis.open_closed(self.Etot)

    def partial_cross_sections(self):
        """Per energy, sigma^{Jp}(a -> b) between the open monomer levels (bohr^2)."""
        from pyticc_amd.crosssection import partial_cross_sections

        return partial_cross_sections(self)

    @property
    def Y_asymptotic(self) -> np.ndarray:
        return np.einsum("ia,...ij,jb->...ab", self.asymptotic_transform, self.Y_propagated, self.asymptotic_transform, optimize=True)

    @property
    def open_channel_indices(self) -> tuple:
        return tuple(np.asarray(np.flatnonzero(m), dtype=np.int64) for m in self.open_closed.open_mask)


@dataclass(frozen=True)
class KBlockResult:
    """Matched result of one CS or NNCC block (result.py:146-205).

    open_channel_indices[i]: global positions of the block's channels open at
    energy i; Y_BF (n_E, n_b, n_b); Bmat, L of the block; Smat_asymptotic[i]
    (n_open_i, n_open_i) in the block's asymptotic basis.
    """

    block: object
    open_channel_indices: tuple
    Y_BF: np.ndarray
    Bmat: np.ndarray
    L: np.ndarray
    Smat_asymptotic: tuple

    @property
    def Y_asymptotic(self) -> np.ndarray:
        return np.einsum("ia,...ij,jb->...ab", self.Bmat, self.Y_BF, self.Bmat, optimize=True)

    @property
    def Smat_BF(self) -> tuple:
        """S_BF = B_open S_asym B_open^T, labelled by the complete-basis K channels."""
        local = {g: i for i, g in enumerate(self.block.channel_indices)}
        out = []
        for open_idx, S in zip(self.open_channel_indices, self.Smat_asymptotic):
            pos = np.asarray([local[int(g)] for g in open_idx], dtype=np.int64)
            B_open = self.Bmat[np.ix_(pos, pos)]
            out.append(np.asarray(B_open @ S @ B_open.T, dtype=np.complex128))
        return tuple(out)


@dataclass(frozen=True)
class CoupledStatesResult:
    """CS or NNCC result of one (J, parity) block: independently matched K blocks (result.py:207-238)."""

    basis: object
    Etot: np.ndarray
    approx: object
    blocks: tuple
    timing: Timing | None = None

    @property
    def molecule_exchange(self) -> int:
        return self.basis.molecule_exchange

    @property
    def open_closed(self):
        return self.basis.open_closed(self.Etot)
