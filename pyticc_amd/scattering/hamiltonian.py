"""Projected channel Hamiltonian (src/pyticc/scattering/hamiltonian.py:36-179 and the
build_hamiltonian of src/pyticc/scattering/energy_transfer/{atom_diatom,diatom_diatom}.py).

W_points is the device seam that replaces V_blocks_device + the W assembly:
it contracts the cached device PES at the requested radial points straight
into W_base = U/R^2 + 2 mu V + 2 mu diag(E_int).
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace
from math import prod

import numpy as np
import torch

from pyticc_amd import _device
from pyticc_amd._log import fail
from pyticc_amd.matrix.centrifugal import get_Umat_BF
from pyticc_amd.matrix.interaction import VBasisBF, contract_w_device, prepare_atom_diatom, prepare_diatom_diatom
from pyticc_amd.system import Approx, ScatteringType


@dataclass(frozen=True)
class ScattHamiltonian:
    basis: object
    reduced_mass: float
    vbasis: VBasisBF
    potential_grid: object
    approx: Approx = Approx.EXACT
    K_delta: int = 1
    _cache: dict = field(default_factory=dict, compare=False, repr=False)
    # set by truncate(): the untruncated Hamiltonian this one is the leading block of
    truncated_from: object = field(default=None, compare=False, repr=False)

    @property
    def E_int(self) -> np.ndarray:
        return self.basis.E_int

    @property
    def potential_grid_size(self) -> int:
        return prod(self.vbasis.grid_shape)

    def centrifugal(self, channel_indices=None) -> np.ndarray:
        return get_Umat_BF(self.basis, channel_indices)

    @property
    def U(self) -> np.ndarray:
        return self.centrifugal()

    def _device_consts(self, device: torch.device):
        key = (device.type, device.index)
        c = self._cache.get(key)
        if c is None:
            c = (_device.to_device(self.U, device), _device.to_device(self.E_int, device))
            self._cache[key] = c
        return c

    def block(self, channel_indices) -> "ScattHamiltonian":
        """The Hamiltonian projected on a CS/NNCC block (whole K subsets): its
        channels, their weighted basis rows, U restricted to them (Coriolis
        couplings only between K values inside the block)."""
        indices = tuple(int(i) for i in channel_indices)
        basis = replace(self.basis, channels=tuple(self.basis[i] for i in indices))
        return ScattHamiltonian(basis, self.reduced_mass, self.vbasis.restrict(indices), self.potential_grid, self.approx,
                                self.K_delta)

    def truncate(self, n_channel: int) -> "ScattHamiltonian":
        """The Hamiltonian on the n_channel lowest channels (the basis is sorted
        by threshold energy, channel.py:569): W_base of the truncated basis is
        the leading n_channel x n_channel block of the full one."""
        n = int(n_channel)
        if not 1 <= n <= self.basis.n_channel:
            raise fail(ValueError, f"cannot truncate {self.basis.n_channel} channels to {n}")
        indices = tuple(range(n))
        basis = replace(self.basis, channels=tuple(self.basis[i] for i in indices))
        root = self if self.truncated_from is None else self.truncated_from
        return ScattHamiltonian(basis, self.reduced_mass, self.vbasis.restrict(indices, whole_k=False),
                                self.potential_grid, self.approx, self.K_delta, truncated_from=root)

    def W_points(self, radial_points, device: torch.device, out: torch.Tensor | None = None) -> torch.Tensor:
        """Energy-independent W_base at the given grid points, (n_R, N, N) on the device."""
        U, Ei = self._device_consts(device)
        V = self.potential_grid.take_device(np.asarray(radial_points), device)
        return contract_w_device(self.vbasis, radial_points, V, U, Ei, self.reduced_mass, device, out=out)


def build_hamiltonian(system, potential_grid) -> ScattHamiltonian:
    """Weighted rows B_K from the grid's quadrature (energy_transfer/*.py build_hamiltonian)."""
    if system.reduced_mass is None:
        raise fail(ValueError, "The Hamiltonian requires a collision reduced mass")
    if potential_grid.scattering_type is not system.scattering_type:
        raise fail(TypeError, f"System type {system.scattering_type.value} cannot use a {potential_grid.scattering_type.value} potential grid")
    if system.scattering_type is ScatteringType.ATOM_DIATOM:
        vb = prepare_atom_diatom(system.basis, system.monomer_Y.rovib, potential_grid.coordinate("cos_theta"), potential_grid.weight("theta"))
    else:
        vb = prepare_diatom_diatom(
            system.basis,
            system.monomer_X.rovib,
            system.monomer_Y.rovib,
            potential_grid.coordinate("cos_theta_X"),
            potential_grid.weight("theta_X"),
            potential_grid.coordinate("cos_theta_Y"),
            potential_grid.weight("theta_Y"),
            potential_grid.coordinate("phi"),
            potential_grid.weight("phi"),
        )
    return ScattHamiltonian(system.basis, system.reduced_mass, vb, potential_grid, system.approx, system.K_delta)
