"""``run(source, *, pes=None)``: a scattering calculation from a TOML input file
(src/pyticc/input/driver.py:147-204, input/atom_diatom.py:126-170,
input/diatom_diatom.py:25-72).

The two geometries of this backend are read: "A+BC" (keys atom, diatom, Jtot,
system_parity, energies_cm and the sections basis, quadrature, channels,
approximation, potential_grid, propagation) and "AB+CD" (diatom_X, diatom_Y,
basis_X, basis_Y, ...), with the reference's key names and units.  The PES comes
from ``pes=`` or from a ``[pes] library = "..."`` entry naming a compiled surface
with the C entry points of include/ticc_pes.h (the reference compiles Fortran
sources there, which this backend does not).  External-field, diabatic and
fine-structure inputs are recognised and refused.
"""

from __future__ import annotations

import sys
from pathlib import Path
from time import perf_counter, process_time

if sys.version_info >= (3, 11):
    import tomllib
else:  # the same parser, packaged separately before Python 3.11
    import tomli as tomllib

from pyticc_amd._log import fail, logger
from pyticc_amd.constants import CM2AU
from pyticc_amd.input.common import (
    approximation,
    build_diatom,
    diatom_symbols,
    energies,
    k_cut,
    potential_grid_settings,
    propagation,
    required,
    resolve_path,
    section,
)
from pyticc_amd.pes import PESWrapper
from pyticc_amd.result import Timing
from pyticc_amd.system import ChannelSpec, ScatteringType, build_ScattSystem, element_mass_au, reduced_mass

_REFUSED_TYPES = ("A+B", "A+BC_electric", "A+BC_diabatic", "A+BC_both_fs", "AB+CD_fs")


def _resolve_scattering_type(config: dict) -> ScatteringType:
    """The calculation path from `type` and the geometry sections (driver.py:43-144)."""
    type_value = str(config.get("type", "auto"))
    if "fine_structure" in config:
        raise fail(NotImplementedError, "Fine-structure inputs are not supported by the MI355X backend")
    if "electric" in config:
        raise fail(NotImplementedError, "External-field inputs are not supported by the MI355X backend")
    if type_value == "auto":
        if "atom" in config and "diatom" in config:
            return ScatteringType.ATOM_DIATOM
        if "diatom_X" in config and "diatom_Y" in config:
            return ScatteringType.DIATOM_DIATOM
        raise fail(ValueError, "Cannot infer TOML scattering geometry; provide atom/diatom or diatom_X/diatom_Y")
    if type_value in _REFUSED_TYPES:
        raise fail(NotImplementedError, f"Calculation type {type_value!r} is not supported by the MI355X backend")
    try:
        return ScatteringType(type_value)
    except ValueError as error:
        supported = ", ".join(t.value for t in ScatteringType)
        raise fail(ValueError, f"Unsupported TOML type {type_value!r}; supported: auto, {supported}") from error


def _load_pes(config: dict, base: Path) -> PESWrapper:
    from pyticc_amd.pes.native import load_native_pes

    values = section(config, "pes")
    if "library" not in values:
        raise fail(NotImplementedError, "The [pes] section must name a compiled C-ABI surface as library = \"...\" "
                                        "(Fortran PES sources are not compiled by this backend), or pass pes=...")
    return load_native_pes(resolve_path(base, values["library"]), monomer_X=values.get("monomer_X"),
                           monomer_Y=values.get("monomer_Y"))


def _channel_spec(channels: dict, two_diatoms: bool) -> ChannelSpec:
    common = dict(vmin_Y=int(channels.get("vmin_Y", 0)), exchange_parity_Y=int(channels.get("exchange_parity_Y", 0)),
                  E_Y_cut=float(required(channels, "E_Y_cut_cm")) * CM2AU, K_cut=k_cut(channels))
    if two_diatoms:
        common.update(vmin_X=int(channels.get("vmin_X", 0)), exchange_parity_X=int(channels.get("exchange_parity_X", 0)),
                      E_X_cut=float(required(channels, "E_X_cut_cm")) * CM2AU)
    return ChannelSpec(**common)


def build_system(config: dict, scattering_type: ScatteringType, pes: PESWrapper):
    """(ScattSystem, quadrature options) of a parsed input."""
    from pyticc_amd.basis.monomer import AtomSpec

    approx, K_delta = approximation(config)
    channels = section(config, "channels")
    quadrature = section(config, "quadrature")
    common = dict(Jtot=int(required(config, "Jtot")), system_parity=int(required(config, "system_parity")),
                  approx=approx, K_delta=K_delta, potential=pes)
    if scattering_type is ScatteringType.ATOM_DIATOM:
        if pes.monomer_Y is None:
            raise fail(ValueError, "Atom-diatom calculation requires the diatomic monomer potential as pes.monomer_Y")
        diatom, m_diatom = build_diatom(diatom_symbols(config, "diatom"), section(config, "basis"), pes.monomer_Y)
        m_atom = element_mass_au(str(required(config, "atom")))
        system = build_ScattSystem(AtomSpec(), diatom, channel=_channel_spec(channels, False),
                                   reduced_mass=reduced_mass(m_atom, m_diatom), **common)
        return system, {"n_theta": int(required(quadrature, "n_theta"))}
    if config.get("molecule_exchange", 0) != 0:
        raise fail(NotImplementedError, "molecule_exchange requires the Python API with a shared monomer basis object")
    if pes.monomer_X is None or pes.monomer_Y is None:
        raise fail(ValueError, "Diatom-diatom calculation requires both pes.monomer_X and pes.monomer_Y")
    basis_X, m_X = build_diatom(diatom_symbols(config, "diatom_X"), section(config, "basis_X"), pes.monomer_X)
    basis_Y, m_Y = build_diatom(diatom_symbols(config, "diatom_Y"), section(config, "basis_Y"), pes.monomer_Y)
    system = build_ScattSystem(basis_X, basis_Y, channel=_channel_spec(channels, True),
                               reduced_mass=reduced_mass(m_X, m_Y), **common)
    quad = {k: int(required(quadrature, k)) for k in ("n_theta_X", "n_theta_Y", "n_phi")}
    return system, quad


def run(source, *, pes: PESWrapper | None = None):
    """Run a scattering calculation from a TOML input file (driver.py:147-204)."""
    from pyticc_amd.scattering import prepare_potential, solve

    wall0, cpu0 = perf_counter(), process_time()
    path = Path(source).expanduser().resolve()
    try:
        with path.open("rb") as fh:
            config = tomllib.load(fh)
    except (OSError, tomllib.TOMLDecodeError) as error:
        raise fail(ValueError, f"Failed to read PyTICC input {path}: {error}") from error
    scattering_type = _resolve_scattering_type(config)
    potential = _load_pes(config, path.parent) if pes is None else pes
    if not isinstance(potential, PESWrapper):
        raise fail(TypeError, f"Calculation type {scattering_type.value!r} requires PESWrapper")
    system, quad = build_system(config, scattering_type, potential)
    boundaries, half_steps, processes = potential_grid_settings(config)
    grid = prepare_potential(system, boundaries, half_steps, processes=processes, **quad)
    result = solve(system, energies(required(config, "energies_cm"), path.parent), grid, propagation(config))
    total = Timing(wall_seconds=perf_counter() - wall0, cpu_seconds=process_time() - cpu0)
    logger.info(f"Calculation complete: total {total}")
    return result
