"""PES callback ABI and tensor-grid evaluation (src/pyticc/pes/adiabatic.py:17-137, 224-256)."""

from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass, field

import numpy as np

from pyticc_amd._log import fail


@dataclass(frozen=True)
class PESWrapper:
    """interaction(R, coords (n_coord, n_grid)) -> (n_grid,); optional interaction_many(R (n_R,), coords) -> (n_R, n_grid)."""

    interaction: Callable
    monomer_X: Callable | None = None
    monomer_Y: Callable | None = None
    interaction_many: Callable | None = None
    _interaction_many_processes: Callable | None = field(default=None, repr=False, compare=False)
    _close: Callable | None = field(default=None, repr=False, compare=False)

    def close(self) -> None:
        if self._close is not None:
            self._close()


def _evaluate(pes: PESWrapper, R, coordinates: np.ndarray, grid_shape, *, processes: int) -> np.ndarray:
    if not isinstance(processes, int) or isinstance(processes, bool) or processes < 1:
        raise fail(ValueError, f"processes must be a positive integer, but got {processes!r}")
    Rs = np.asarray(R, dtype=np.float64)
    if Rs.ndim == 0:
        values = np.asarray(pes.interaction(float(Rs), coordinates), dtype=np.float64)
        expected = (coordinates.shape[1],)
        out_shape = tuple(grid_shape)
    elif Rs.ndim == 1:
        if Rs.size == 0:
            values = np.empty((0, coordinates.shape[1]))
        elif pes._interaction_many_processes is not None:
            values = np.asarray(pes._interaction_many_processes(Rs, coordinates, processes), dtype=np.float64)
        elif pes.interaction_many is None:
            values = np.stack([pes.interaction(float(r), coordinates) for r in Rs])
        else:
            values = np.asarray(pes.interaction_many(Rs, coordinates), dtype=np.float64)
        expected = (Rs.size, coordinates.shape[1])
        out_shape = (Rs.size, *grid_shape)
    else:
        raise fail(ValueError, f"R must be a scalar or one-dimensional array, but got shape {Rs.shape}")
    if values.shape != expected:
        raise fail(ValueError, f"Interaction PES returned shape {values.shape}, but expected {expected}")
    return values.reshape(out_shape)


def get_Vgrid_atom_diatom(pes, R, r, theta, *, processes: int = 1) -> np.ndarray:
    """Coordinates rows (r, theta), Fortran order; returns (n_r, n_theta) or (n_R, n_r, n_theta)."""
    grids = np.meshgrid(r, theta, indexing="ij")
    coords = np.asfortranarray(np.stack([g.reshape(-1) for g in grids]))
    return _evaluate(pes, R, coords, grids[0].shape, processes=processes)


def get_Vgrid_diatom_diatom(pes, R, r_X, r_Y, theta_X, theta_Y, phi, *, processes: int = 1) -> np.ndarray:
    """Coordinates rows (r_X, r_Y, theta_X, theta_Y, phi), Fortran order."""
    grids = np.meshgrid(r_X, r_Y, theta_X, theta_Y, phi, indexing="ij")
    coords = np.asfortranarray(np.stack([g.reshape(-1) for g in grids]))
    return _evaluate(pes, R, coords, grids[0].shape, processes=processes)
