"""Monomer models of the closed-shell exact-CC path.

AtomSpec restates src/pyticc/basis/monomer/atom.py:1-24; DiatomSpec and
DiatomBasis restate src/pyticc/basis/monomer/diatom.py:37-139.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from pyticc_amd._log import fail
from pyticc_amd.basis.rovib import RovibBasis
from pyticc_amd.system import MolInnerState, MonomerType


class AtomSpec:
    """Structureless atom with one zero-energy internal state."""

    type = MonomerType.ATOM

    def mis_iter(self, E_cut: float):
        yield MolInnerState(j=0, Eint=0.0)

    def energy(self, mis: MolInnerState, K: int) -> float:
        return 0.0

    def allows_K(self, mis: MolInnerState, K: int) -> bool:
        return True


def _validate_state_table(Eint: np.ndarray) -> None:
    if Eint.ndim != 2:
        raise fail(ValueError, f"Eint must be a two-dimensional array indexed as Eint[v, j], but got ndim={Eint.ndim}")
    if 0 in Eint.shape:
        raise fail(ValueError, f"Eint must contain at least one vibrational and rotational level, but got shape {Eint.shape}")


@dataclass(frozen=True)
class DiatomSpec:
    """Lightweight diatomic state table Eint[v, j] used for channel enumeration."""

    Eint: np.ndarray
    type = MonomerType.DIATOM

    def __post_init__(self) -> None:
        _validate_state_table(np.asarray(self.Eint))

    @property
    def vmax(self) -> int:
        return self.Eint.shape[0] - 1

    @property
    def jmax(self) -> int:
        return self.Eint.shape[1] - 1

    def mis_iter(self, E_cut: float):
        for v in range(self.vmax + 1):
            for j in range(self.jmax + 1):
                energy = float(self.Eint[v, j])
                if np.isfinite(energy) and energy <= E_cut:
                    yield MolInnerState(j=j, v=v, Eint=energy)

    def energy(self, mis: MolInnerState, K: int) -> float:
        if mis.v is None:
            raise fail(ValueError, "Diatomic inner state requires v")
        return float(self.Eint[mis.v, mis.j])

    def allows_K(self, mis: MolInnerState, K: int) -> bool:
        return True


@dataclass(frozen=True)
class DiatomBasis:
    """Adiabatic diatom: rovibrational basis and the energy zero of channel thresholds."""

    rovib: RovibBasis
    energy_zero: float
    type = MonomerType.DIATOM

    def __post_init__(self) -> None:
        if not np.isfinite(self.energy_zero):
            raise fail(ValueError, f"energy_zero must be finite, but got {self.energy_zero}")
        _validate_state_table(np.asarray(self.rovib.E_vj))

    @property
    def vmax(self) -> int:
        return self.rovib.E_vj.shape[0] - 1

    @property
    def jmax(self) -> int:
        return self.rovib.E_vj.shape[1] - 1

    @property
    def Eint(self) -> np.ndarray:
        return np.asarray(self.rovib.E_vj - self.energy_zero, dtype=np.float64)

    def mis_iter(self, E_cut: float):
        for v in range(self.vmax + 1):
            for j in range(self.jmax + 1):
                energy = float(self.rovib.E_vj[v, j] - self.energy_zero)
                if np.isfinite(energy) and energy <= E_cut:
                    yield MolInnerState(j=j, v=v, Eint=energy)

    def energy(self, mis: MolInnerState, K: int) -> float:
        if mis.v is None:
            raise fail(ValueError, "Diatomic inner state requires v")
        return float(self.rovib.E_vj[mis.v, mis.j] - self.energy_zero)

    def allows_K(self, mis: MolInnerState, K: int) -> bool:
        return True


def rigid_rotor(B_rot: float, jmax: int, r_e: float = 1.0) -> DiatomBasis:
    """Rigid rotor: E_j = B j(j+1), one radial grid point with WF = 1 (the reference tests' fixture)."""
    j = np.arange(jmax + 1)
    rovib = RovibBasis(grids=np.array([float(r_e)]), E_vj=(B_rot * j * (j + 1.0))[None, :], WF_vj=np.ones((1, 1, jmax + 1)))
    return DiatomBasis(rovib=rovib, energy_zero=0.0)


def build_DiatomBasis(dvr, *, n_podvr: int, vmax: int, jmax: int, mass: float,
                      energy_zero: float | None = None) -> DiatomBasis:
    """Adiabatic diatom from a primitive sine DVR through the PODVR contraction
    (basis/monomer/diatom.py:141-158); the energy zero defaults to E(v=0, j=0)."""
    from pyticc_amd.basis.dvr import build_RovibPODVR

    rovib = build_RovibPODVR(dvr, n_podvr, vmax, jmax, mass)
    zero = float(rovib.E_vj[0, 0]) if energy_zero is None else float(energy_zero)
    return DiatomBasis(rovib=rovib, energy_zero=zero)


def prepare_Diatom(potential, *, r: tuple, n_dvr: int, n_podvr: int, vmax: int, jmax: int, mass: float,
                   energy_zero: float | None = None) -> DiatomBasis:
    """Diatomic monomer from its potential: sine DVR on r = (a, b), then PODVR
    (basis/monomer/diatom.py:162-207)."""
    from pyticc_amd.basis.dvr import build_SineDVR

    if potential is None:
        raise fail(ValueError, "Diatomic monomer preparation requires a monomer potential")
    dvr = build_SineDVR(float(r[0]), float(r[1]), int(n_dvr), float(mass), potential)
    return build_DiatomBasis(dvr, n_podvr=n_podvr, vmax=vmax, jmax=jmax, mass=mass, energy_zero=energy_zero)
