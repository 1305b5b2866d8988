"""ticc.solve for the exact coupled-channel path (src/pyticc/scattering/solver.py:57-196).

Everything after the one-time host setup (channel list, weighted basis rows,
U, BF->SF eigenvectors) runs on the device: contraction, LDMD sweep and the
fused BF->SF + matching kernel.
"""

from __future__ import annotations

from dataclasses import replace
from time import perf_counter, process_time

import numpy as np

from pyticc_amd import _device
from pyticc_amd._log import fail, logger
from pyticc_amd.energy import get_Etot
from pyticc_amd.match.asymptotic import get_Bmat_BF_to_SF
from pyticc_amd.match.smatrix import get_Smat_device, unpack_smat
from pyticc_amd.propagation.device import resolve_device
from pyticc_amd.propagation.runner import propagate
from pyticc_amd.result import ScatteringResult, Timing
from pyticc_amd.scattering.hamiltonian import ScattHamiltonian, build_hamiltonian
from pyticc_amd.scattering.potential import PotentialGrid
from pyticc_amd.system import Approx, ScattSystem


def solve(problem, Etot, potential_grid, propagation) -> ScatteringResult:
    """Build, propagate and match one exact-CC block; returns ScatteringResult with timing."""
    if isinstance(problem, ScattSystem):
        if not isinstance(potential_grid, PotentialGrid):
            raise fail(TypeError, "A ScattSystem solve requires a PotentialGrid")
        if problem.scattering_type is None:
            raise fail(TypeError, "A ScattSystem solve requires a system built by build_ScattSystem")
        hamiltonian = build_hamiltonian(problem, potential_grid)
        sectors = potential_grid.sectors
    elif isinstance(problem, ScattHamiltonian):
        hamiltonian = problem
        sectors = tuple(potential_grid) if not isinstance(potential_grid, PotentialGrid) else potential_grid.sectors
    else:
        raise fail(TypeError, f"solve expects a ScattSystem or ScattHamiltonian, got {type(problem).__name__}")
    if hamiltonian.approx is not Approx.EXACT:
        raise fail(NotImplementedError, "CS/NNCC block propagation is not supported by the MI355X backend")
    energies = get_Etot(Etot)
    basis = hamiltonian.basis
    wall0, cpu0 = perf_counter(), process_time()
    logger.info(f"Solving Jtot={basis.Jtot}, system_parity={basis.system_parity:+d}, approx=exact, channels={basis.n_channel}, energies={energies.size}")
    B, L = get_Bmat_BF_to_SF(basis)
    Y, S_re, S_im, n_open = solve_device(hamiltonian, energies, sectors, propagation, B, L)
    Smat = unpack_smat(S_re, S_im, n_open)
    result = ScatteringResult(basis=basis, Etot=energies, Y_propagated=_device.to_host(Y), asymptotic_transform=B, L=L, Smat=Smat)
    timing = Timing(wall_seconds=perf_counter() - wall0, cpu_seconds=process_time() - cpu0)
    logger.info(f"Solver complete: {timing}")
    return replace(result, timing=timing)


def solve_device(hamiltonian, energies, sectors, propagation, B, L):
    """Device hot path: contraction + LDMD sweep + matching; returns device (Y, S_re, S_im) and n_open."""
    if propagation.mode != "inelastic":
        raise fail(NotImplementedError, "capture-mode propagation is not supported by the MI355X backend")
    Y = propagate(hamiltonian, energies, sectors, propagation)
    dev = resolve_device(propagation.device).torch_device
    S_re, S_im, n_open = get_Smat_device(Y, B, sectors[-1].radial_end, energies, hamiltonian.reduced_mass, hamiltonian.basis.E_int, L, dev)
    return Y, S_re, S_im, n_open
