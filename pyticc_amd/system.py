"""Problem definition: monomers, selection rules, one (J, parity) block.

Mirrors src/pyticc/system.py of the reference (MolInnerState 152-158,
ChannelSpec 243-299, ScattSystem 306-414, _infer_scattering_type 421-459,
build_ScattSystem 466-777) for the two geometries of the exact
coupled-channel path: atom + diatom (A+BC) and diatom + diatom (AB+CD).
Other reference geometries (fine structure, electric field, diabatic,
Delves, triatom) are rejected with TypeError.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, replace
from enum import Enum

import numpy as np

from pyticc_amd._log import fail
from pyticc_amd.constants import AMU2AU

ELEMENT_MASS_AMU: dict[str, float] = {
    # most abundant isotope (standard atomic weight for He, Li, S), as the reference's table
    "H": 1.00782503223,
    "D": 2.01410177812,
    "B": 11.00930536,
    "He": 4.002602,
    "Li": 6.938,
    "N": 14.00307400443,
    "O": 15.99491461957,
    "C": 12.0,
    "F": 18.99840316273,
    "S": 32.06,
    "Cl": 34.968852682,
    "Ar": 39.9623831237,
    "K": 39.963998166,
    "Rb": 86.909180531,
}


def element_mass_au(symbol: str) -> float:
    """Atomic mass of a supported element in atomic units (system.py:57-74)."""
    if symbol not in ELEMENT_MASS_AMU:
        raise fail(ValueError, f"Unsupported element symbol {symbol!r}. Supported: {', '.join(ELEMENT_MASS_AMU)}")
    return ELEMENT_MASS_AMU[symbol] * AMU2AU


def reduced_mass(mass1: float, mass2: float) -> float:
    """mu = m1 m2 / (m1 + m2) (system.py:96-114)."""
    m1 = float(mass1)
    m2 = float(mass2)
    return m1 * m2 / (m1 + m2)


class MonomerType(Enum):
    ATOM = "1"
    DIATOM = "2"
    TRIATOM = "3"


class Approx(Enum):
    EXACT = "exact"
    CS = "cs"
    NNCC = "nncc"


class ScatteringType(Enum):
    """Scattering geometries; this backend propagates ATOM_DIATOM and DIATOM_DIATOM."""

    ATOM_DIATOM = "A+BC"
    DIATOM_DIATOM = "AB+CD"


@dataclass(frozen=True)
class MolInnerState:
    j: int
    v: int | None = None
    t: int | None = None
    Eint: float = 0.0
    electronic_state: int | None = None


@dataclass(frozen=True, slots=True)
class ChannelSpec:
    """Selection rules and truncations used to construct channels (system.py:243-299)."""

    vmin_X: int | tuple[int, ...] = 0
    vmin_Y: int | tuple[int, ...] = 0
    exchange_parity_X: int | tuple[int, ...] = 0
    exchange_parity_Y: int | tuple[int, ...] = 0
    E_X_cut: float = math.inf
    E_Y_cut: float = math.inf
    K_cut: int | None = None

    def __post_init__(self) -> None:
        for name, values in (("vmin_X", self.vmin_X), ("vmin_Y", self.vmin_Y)):
            sel = (values,) if isinstance(values, int) else values
            if not sel or any(v < 0 for v in sel):
                raise fail(ValueError, f"{name} must contain non-negative integers, but got {values!r}")
        for name, values in (("exchange_parity_X", self.exchange_parity_X), ("exchange_parity_Y", self.exchange_parity_Y)):
            sel = (values,) if isinstance(values, int) else values
            if not sel or any(v not in (-1, 0, 1) for v in sel):
                raise fail(ValueError, f"{name} must contain only -1, 0, or 1, but got {values!r}")
        if any(np.isnan(v) or np.isneginf(v) for v in (self.E_X_cut, self.E_Y_cut)):
            raise fail(ValueError, f"Energy cutoffs must be finite or positive infinity, but got E_X_cut={self.E_X_cut}, E_Y_cut={self.E_Y_cut}")
        if self.K_cut is not None and self.K_cut < 0:
            raise fail(ValueError, f"K_cut must be non-negative, but got K_cut={self.K_cut}")


@dataclass(frozen=True)
class ScattSystem:
    """Physical definition of one scattering block (system.py:306-414)."""

    monomer_X: object
    monomer_Y: object = None
    scattering_type: ScatteringType | None = None
    Jtot: int | None = None
    two_J: int | None = None
    system_parity: int | None = None
    M: int | None = None
    approx: Approx = Approx.EXACT
    K_delta: int = 1
    potential: object = None
    total_potential: object = None
    reduced_mass: float | None = None
    magnetic_dipole_coefficient: float = 0.0
    channel_spec: ChannelSpec | None = None
    basis: object = None
    molecule_exchange: int = 0

    def __post_init__(self) -> None:
        if isinstance(self.molecule_exchange, bool) or self.molecule_exchange not in (-1, 0, 1):
            raise fail(ValueError, "molecule_exchange must be -1, 0, or 1")
        if self.Jtot is not None and self.Jtot < 0:
            raise fail(ValueError, f"Invalid Jtot {self.Jtot}. Must be non-negative.")
        if self.system_parity is not None and self.system_parity not in (-1, 1):
            raise fail(ValueError, f"Invalid system_parity {self.system_parity}. Must be -1 or 1.")
        if self.K_delta < 1:
            raise fail(ValueError, f"Invalid K_delta {self.K_delta}. Must be positive.")
        if self.reduced_mass is not None and self.reduced_mass <= 0.0:
            raise fail(ValueError, f"Invalid reduced_mass {self.reduced_mass}. Must be positive.")

    @property
    def n_channel(self) -> int:
        if self.basis is None:
            raise fail(ValueError, "Scattering channels have not been prepared; use build_ScattSystem")
        return self.basis.n_channel


def _infer_scattering_type(monomer_X, monomer_Y) -> ScatteringType:
    """Geometry from the monomer types (system.py:421-459, closed-shell rows)."""
    from pyticc_amd.basis.monomer import AtomSpec

    tx = getattr(monomer_X, "type", None)
    ty = getattr(monomer_Y, "type", None)
    if isinstance(monomer_X, AtomSpec) and ty is MonomerType.DIATOM:
        return ScatteringType.ATOM_DIATOM
    if tx is MonomerType.DIATOM and ty is MonomerType.DIATOM:
        return ScatteringType.DIATOM_DIATOM
    raise fail(
        TypeError,
        f"Unsupported scattering monomer combination: monomer_X={type(monomer_X).__name__}, monomer_Y={type(monomer_Y).__name__}",
    )


def build_ScattSystem(
    monomer_X,
    monomer_Y=None,
    *,
    Jtot: int | None = None,
    system_parity: int | None = None,
    channel: ChannelSpec | None = None,
    approx: Approx = Approx.EXACT,
    K_delta: int = 1,
    potential=None,
    reduced_mass: float | None = None,
    molecule_exchange: int = 0,
) -> ScattSystem:
    """Build a scattering system and its channel basis (system.py:466-777)."""
    from pyticc_amd.basis.channel import build_ChannelBasis

    if monomer_Y is None:
        raise fail(ValueError, "Nonreactive system construction requires two prepared monomers")
    selected = _infer_scattering_type(monomer_X, monomer_Y)
    if molecule_exchange:
        raise fail(NotImplementedError, "molecule_exchange channels are not supported by the MI355X backend")
    spec = ChannelSpec() if channel is None else channel
    system = ScattSystem(
        monomer_X=monomer_X,
        monomer_Y=monomer_Y,
        scattering_type=selected,
        Jtot=Jtot,
        system_parity=system_parity,
        approx=approx,
        K_delta=K_delta,
        potential=potential,
        reduced_mass=reduced_mass,
        channel_spec=spec,
        molecule_exchange=molecule_exchange,
    )
    if Jtot is None or system_parity is None:
        raise fail(ValueError, "Field-free system construction requires Jtot and system_parity")
    return replace(system, basis=build_ChannelBasis(system, spec))
