"""Public solve / prepare_potential surface of the exact-CC path."""

from pyticc_amd.scattering.hamiltonian import ScattHamiltonian
from pyticc_amd.scattering.potential import PotentialGrid, prepare_potential
from pyticc_amd.scattering.solver import solve

__all__ = ["PotentialGrid", "ScattHamiltonian", "prepare_potential", "solve"]
