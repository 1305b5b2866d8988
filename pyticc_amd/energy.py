"""Total-energy input (src/pyticc/energy.py:12-55 of the reference)."""

from __future__ import annotations

from pathlib import Path

import numpy as np

from pyticc_amd._log import fail


def get_Etot(source) -> np.ndarray:
    """Read-only FP64 energies from an array-like or a one-column text file."""
    if isinstance(source, (str, Path)):
        try:
            energies = np.loadtxt(Path(source), dtype=np.float64, comments="#", ndmin=1)
        except (OSError, ValueError) as error:
            raise fail(ValueError, f"Failed to read total energies from {source}: {error}") from error
    else:
        try:
            energies = np.asarray(source, dtype=np.float64)
        except (TypeError, ValueError) as error:
            raise fail(ValueError, f"Failed to convert total energies to an array: {error}") from error
    if energies.ndim != 1:
        raise fail(ValueError, f"Total energies must be one-dimensional, but got shape={energies.shape}")
    if energies.size == 0:
        raise fail(ValueError, "At least one total energy is required")
    if not np.all(np.isfinite(energies)):
        raise fail(ValueError, "Total energies must be finite")
    energies = energies.copy()
    energies.setflags(write=False)
    return energies
