"""Raw PES on every propagation point (src/pyticc/scattering/potential.py:26-242,
src/pyticc/scattering/energy_transfer/{atom_diatom,diatom_diatom}.py prepare_potential)."""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass, field
from time import perf_counter

import numpy as np
import torch

from pyticc_amd._log import fail, logger
from pyticc_amd.basis.angle import gauss_legendre_dvr
from pyticc_amd.pes import PESWrapper, get_Vgrid_atom_diatom, get_Vgrid_diatom_diatom
from pyticc_amd.propagation.grid import RadialSector, build_radial_sectors
from pyticc_amd.system import ScatteringType


@dataclass(frozen=True, slots=True)
class PotentialGrid:
    """Raw interaction values with the ordered radial points as the leading axis; one HBM copy per device."""

    boundaries: tuple
    half_steps: tuple
    sectors: tuple
    radial_points: np.ndarray
    scattering_type: ScatteringType
    coordinates: tuple
    weights: tuple
    values: np.ndarray
    _device_cache: dict = field(default_factory=dict, init=False, repr=False, compare=False)

    @property
    def Rmatch(self) -> float:
        return self.sectors[-1].radial_end

    def coordinate(self, name: str) -> np.ndarray:
        return self._named(self.coordinates, name, "coordinate")

    def weight(self, name: str) -> np.ndarray:
        return self._named(self.weights, name, "weight")

    def take(self, radial_points):
        return self.values[self._selection(radial_points)]

    def take_device(self, radial_points, device: torch.device) -> torch.Tensor:
        """Device slice of the cached grid (the whole grid is copied once per device)."""
        key = (device.type, device.index)
        vals = self._device_cache.get(key)
        if vals is None:
            vals = torch.from_numpy(np.ascontiguousarray(self.values, dtype=np.float64)).to(device)
            self._device_cache[key] = vals
            logger.info(f"Potential grid resident on gpu:{device.index}, size={self.values.nbytes / 1024**2:.1f} MiB")
        sel = self._selection(radial_points)
        if isinstance(sel, int):
            return vals[sel]
        idx = np.asarray(sel)
        if idx.size and np.all(np.diff(idx) == 1):
            return vals[int(idx[0]) : int(idx[-1]) + 1]
        return vals[torch.as_tensor(idx, device=device)]

    def _selection(self, radial_points):
        req = np.asarray(radial_points, dtype=np.float64)
        if req.ndim > 1:
            raise fail(ValueError, f"radial_points must be scalar or one-dimensional, but got shape {req.shape}")
        flat = req.reshape(-1)
        pos = np.searchsorted(self.radial_points, flat)
        inside = pos < self.radial_points.size
        ok = np.zeros(flat.size, dtype=bool)
        ok[inside] = np.isclose(self.radial_points[pos[inside]], flat[inside], rtol=0.0, atol=1.0e-12)
        if not np.all(ok):
            raise fail(ValueError, f"Potential grid does not contain radial points {flat[~ok].tolist()}")
        return int(pos[0]) if req.ndim == 0 else pos.astype(np.int64)

    @staticmethod
    def _named(arrays, name, label):
        for n, v in arrays:
            if n == name:
                return v
        raise fail(KeyError, f"Potential grid has no {label} named {name!r}")


def _radial_grid(boundaries, half_steps):
    b = tuple(float(v) for v in boundaries)
    h = tuple(float(v) for v in half_steps)
    sectors = build_radial_sectors(b, h)
    pts = np.empty(2 * len(sectors) + 1)
    pts[0] = sectors[0].radial_start
    for i, s in enumerate(sectors):
        pts[2 * i + 1] = s.radial_mid
        pts[2 * i + 2] = s.radial_end
    return b, h, sectors, pts


def _prepare_atom_diatom(system, boundaries, half_steps, *, n_theta: int = 16, processes: int = 1):
    from pyticc_amd.basis.monomer import AtomSpec, DiatomBasis

    if not isinstance(system.monomer_X, AtomSpec) or not isinstance(system.monomer_Y, DiatomBasis):
        raise fail(TypeError, "Atom-diatom potential preparation requires AtomSpec and DiatomBasis monomers")
    if not isinstance(system.potential, PESWrapper):
        raise fail(TypeError, "Atom-diatom potential preparation requires a scalar PES and prepared channels")
    par = system.basis.channel_spec.exchange_parity_Y
    pars = (par,) if isinstance(par, int) else par
    ct, wt = gauss_legendre_dvr(-1.0, 1.0, n_theta, symmetry=all(p != 0 for p in pars))
    b, h, sectors, pts = _radial_grid(boundaries, half_steps)
    r = system.monomer_Y.rovib.grids
    values = get_Vgrid_atom_diatom(system.potential, pts, r, np.arccos(ct), processes=processes)
    return PotentialGrid(b, h, sectors, pts, ScatteringType.ATOM_DIATOM, (("r", r), ("cos_theta", ct)), (("theta", wt),), values)


def _prepare_diatom_diatom(system, boundaries, half_steps, *, n_theta_X: int = 15, n_theta_Y: int = 15, n_phi: int = 12, processes: int = 1):
    from pyticc_amd.basis.monomer import DiatomBasis

    if not isinstance(system.monomer_X, DiatomBasis) or not isinstance(system.monomer_Y, DiatomBasis):
        raise fail(TypeError, "Diatom-diatom potential preparation requires two DiatomBasis monomers")
    if not isinstance(system.potential, PESWrapper):
        raise fail(TypeError, "Diatom-diatom potential preparation requires a scalar PES")
    ctX, wX = gauss_legendre_dvr(-1.0, 1.0, n_theta_X)
    ctY, wY = gauss_legendre_dvr(-1.0, 1.0, n_theta_Y)
    phi, wphi = gauss_legendre_dvr(0.0, np.pi, n_phi)
    b, h, sectors, pts = _radial_grid(boundaries, half_steps)
    rX = system.monomer_X.rovib.grids
    rY = system.monomer_Y.rovib.grids
    values = get_Vgrid_diatom_diatom(system.potential, pts, rX, rY, np.arccos(ctX), np.arccos(ctY), phi, processes=processes)
    return PotentialGrid(
        b, h, sectors, pts, ScatteringType.DIATOM_DIATOM,
        (("r_X", rX), ("r_Y", rY), ("cos_theta_X", ctX), ("cos_theta_Y", ctY), ("phi", phi)),
        (("theta_X", wX), ("theta_Y", wY), ("phi", wphi)),
        values,
    )


def prepare_potential(system, boundaries: Sequence[float], half_steps: Sequence[float], *, processes: int = 1, **quadrature) -> PotentialGrid:
    """Evaluate the system's PES on every ordered propagation point (potential.py:182-222)."""
    from pyticc_amd.system import ScattSystem

    if not isinstance(system, ScattSystem):
        raise fail(TypeError, "prepare_potential requires a ScattSystem")
    if system.scattering_type is None:
        raise fail(TypeError, "prepare_potential requires a system built by build_ScattSystem")
    n_pts = 2 * len(build_radial_sectors(boundaries, half_steps)) + 1
    started = perf_counter()
    logger.info(f"Potential preparation started: type={system.scattering_type.value}, radial_points={n_pts}, processes={processes}")
    if system.scattering_type is ScatteringType.ATOM_DIATOM:
        grid = _prepare_atom_diatom(system, boundaries, half_steps, processes=processes, **quadrature)
    else:
        grid = _prepare_diatom_diatom(system, boundaries, half_steps, processes=processes, **quadrature)
    logger.info(f"Potential preparation complete: type={system.scattering_type.value}, radial_points={n_pts}, wall={perf_counter() - started:.3f} s")
    return grid


__all__ = ["PotentialGrid", "RadialSector", "prepare_potential"]
