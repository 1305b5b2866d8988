"""Parsers shared by the TOML calculation paths (src/pyticc/input/common.py:1-194).

Every parser reads one input section with the reference's key names and units
(energies and cuts in cm^-1, lengths in bohr) and raises ValueError naming the
key or section on a missing or malformed value.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from pyticc_amd._log import fail
from pyticc_amd.basis.monomer import prepare_Diatom
from pyticc_amd.constants import CM2AU
from pyticc_amd.energy import get_Etot
from pyticc_amd.propagation import Propagation
from pyticc_amd.system import Approx, element_mass_au, reduced_mass


def required(table: dict, key: str):
    if key not in table:
        raise fail(ValueError, f"Missing required input {key!r}")
    return table[key]


def section(config: dict, name: str) -> dict:
    value = required(config, name)
    if not isinstance(value, dict):
        raise fail(ValueError, f"Input section {name!r} must be a TOML table")
    return value


def resolve_path(base: Path, value) -> Path:
    """A user path relative to the input file's directory."""
    path = Path(value).expanduser()
    return path.resolve() if path.is_absolute() else (base / path).resolve()


def energies(value, base: Path) -> np.ndarray:
    """Total energies in cm^-1 (a list, or a one-column file next to the input) in atomic units."""
    source = resolve_path(base, value) if isinstance(value, (str, Path)) else value
    return get_Etot(source) * CM2AU


def k_cut(values: dict):
    """A non-negative helicity cutoff or the literal "none"."""
    value = required(values, "K_cut")
    if value == "none":
        return None
    if isinstance(value, int) and not isinstance(value, bool) and value >= 0:
        return value
    raise fail(ValueError, f"K_cut must be a non-negative integer or 'none', but got {value!r}")


def approximation(config: dict) -> tuple:
    """(Approx, K_delta) from [approximation]; exact CC when the section is absent."""
    values = config.get("approximation", {"method": "exact"})
    if not isinstance(values, dict):
        raise fail(ValueError, "Input section 'approximation' must be a TOML table")
    method = required(values, "method")
    try:
        approx = Approx(str(method).lower())
    except ValueError as error:
        raise fail(ValueError, f"Approximation method must be 'exact', 'cs', or 'nncc', but got {method!r}") from error
    K_delta = int(values.get("K_delta", 1))
    if approx is Approx.NNCC and K_delta < 1:
        raise fail(ValueError, f"NNCC requires K_delta >= 1, but got K_delta={K_delta}")
    return approx, K_delta


def propagation(config: dict) -> Propagation:
    values = section(config, "propagation")
    return Propagation(
        mode=required(values, "mode"),
        memory_mb=float(values.get("memory_limit_mb", 512.0)),
        device=values.get("device", "auto"),
        print_verbose=bool(values.get("print_verbose", False)),
    )


def potential_grid_settings(config: dict) -> tuple:
    """(radial boundaries, half steps, PES worker count) from [potential_grid]."""
    values = section(config, "potential_grid")
    boundaries = tuple(float(v) for v in required(values, "radial_boundaries"))
    half_steps = tuple(float(v) for v in required(values, "radial_half_steps"))
    return boundaries, half_steps, int(values.get("processes", 1))


def diatom_symbols(config: dict, key: str) -> tuple:
    symbols = tuple(str(s) for s in required(config, key))
    if len(symbols) != 2:
        raise fail(ValueError, f"{key} must contain two element symbols, but got {symbols}")
    return symbols


def build_diatom(symbols: tuple, values: dict, potential) -> tuple:
    """(DiatomBasis, total mass) of one diatom from its [basis] table: sine DVR on
    r = [a, b] with n_dvr points, n_podvr PODVR points, v <= vmax, j <= jmax."""
    interval = tuple(float(v) for v in required(values, "r"))
    if len(interval) != 2:
        raise fail(ValueError, f"basis r must contain the two DVR boundaries, but got {interval}")
    m1, m2 = (element_mass_au(s) for s in symbols)
    basis = prepare_Diatom(potential, r=interval, n_dvr=int(required(values, "n_dvr")),
                           n_podvr=int(required(values, "n_podvr")), vmax=int(required(values, "vmax")),
                           jmax=int(required(values, "jmax")), mass=reduced_mass(m1, m2))
    return basis, m1 + m2
