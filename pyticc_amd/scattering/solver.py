"""ticc.solve for the inelastic coupled-channel path (src/pyticc/scattering/solver.py:57-217).

Exact CC propagates the whole (J, parity) block; CS and NNCC propagate and match
each K block independently (solver.py:154-177, propagation/runner.py:326-374,
match/finalize.py:168-225).  Everything after the one-time host setup (channel
list, weighted basis rows, U, BF->SF eigenvectors) runs on the device:
contraction, LDMD sweep and the fused BF->SF + matching kernel.
"""

from __future__ import annotations

from dataclasses import replace
from time import perf_counter, process_time

import numpy as np

from pyticc_amd import _device
from pyticc_amd._log import fail, logger
from pyticc_amd.energy import get_Etot
from pyticc_amd.match.asymptotic import get_Bmat_BF_to_SF
from pyticc_amd.match.capture import from_real_form, get_Smat_capture_device
from pyticc_amd.match.smatrix import get_Smat_device, unpack_smat
from pyticc_amd.propagation.device import resolve_device
from pyticc_amd.propagation.runner import propagate, propagate_blocks
from pyticc_amd.basis.kblock import build_cs_blocks, build_nncc_blocks
from pyticc_amd.result import CoupledStatesResult, KBlockResult, ScatteringResult, Timing
from pyticc_amd.scattering.hamiltonian import ScattHamiltonian, build_hamiltonian
from pyticc_amd.scattering.potential import PotentialGrid
from pyticc_amd.system import Approx, ScattSystem


def solve(problem, Etot, potential_grid, propagation) -> ScatteringResult:
    """Build, propagate and match one (J, parity) block -- exact CC, or CS / NNCC K
    blocks -- in the inelastic or capture mode; returns the result with timing."""
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
    energies = get_Etot(Etot)
    basis = hamiltonian.basis
    wall0, cpu0 = perf_counter(), process_time()
    if hamiltonian.approx is not Approx.EXACT:
        return _solve_blocks(hamiltonian, energies, sectors, propagation, wall0, cpu0)
    logger.info(f"Solving Jtot={basis.Jtot}, system_parity={basis.system_parity:+d}, approx=exact, channels={basis.n_channel}, energies={energies.size}")
    B, L = get_Bmat_BF_to_SF(basis)
    Y, Smat = solve_device(hamiltonian, energies, sectors, propagation, B, L)
    result = ScatteringResult(basis=basis, Etot=energies, Y_propagated=Y, asymptotic_transform=B, L=L, Smat=Smat,
                              reduced_mass=hamiltonian.reduced_mass)
    timing = Timing(wall_seconds=perf_counter() - wall0, cpu_seconds=process_time() - cpu0)
    logger.info(f"Solver complete: {timing}")
    return replace(result, timing=timing)


def solve_device(hamiltonian, energies, sectors, propagation, B, L):
    """Device hot path: contraction + LDMD sweep + matching.  Returns the host Y_BF
    (float64, complex128 in capture mode) and the ragged tuple of S matrices."""
    Y = propagate(hamiltonian, energies, sectors, propagation)
    return match_device(hamiltonian, Y, energies, sectors, propagation, B, L)


def match_device(hamiltonian, Y, energies, sectors, propagation, B, L):
    """Matching of propagated device log derivatives; returns (host Y_BF, S tuple)."""
    dev = resolve_device(propagation.device).torch_device
    Rmatch, mu, E_int = sectors[-1].radial_end, hamiltonian.reduced_mass, hamiltonian.basis.E_int
    if propagation.mode == "capture":
        Smat = get_Smat_capture_device(Y, B, Rmatch, energies, mu, E_int, L, dev)
        return _device.to_host(from_real_form(Y)), Smat
    S_re, S_im, n_open = get_Smat_device(Y, B, Rmatch, energies, mu, E_int, L, dev)
    return _device.to_host(Y), unpack_smat(S_re, S_im, n_open)


def build_k_blocks(hamiltonian) -> tuple:
    """CS or NNCC propagation blocks of one Hamiltonian (solver.py:203-214)."""
    if hamiltonian.approx is Approx.CS:
        return build_cs_blocks(hamiltonian.basis)
    if hamiltonian.approx is Approx.NNCC:
        return build_nncc_blocks(hamiltonian.basis, hamiltonian.K_delta)
    raise fail(ValueError, f"Coupled-states blocks require approx='cs' or 'nncc', but got {hamiltonian.approx.value!r}")


def _solve_blocks(hamiltonian, energies, sectors, propagation, wall0, cpu0) -> CoupledStatesResult:
    """Propagate and match every K block on the device (finalize.py:168-225 per block)."""
    basis = hamiltonian.basis
    blocks = build_k_blocks(hamiltonian)
    logger.info(f"Solving Jtot={basis.Jtot}, system_parity={basis.system_parity:+d}, approx={hamiltonian.approx.value}, "
                f"channels={basis.n_channel}, blocks={len(blocks)}, energies={energies.size}")
    subs = [hamiltonian.block(block.channel_indices) for block in blocks]
    # inelastic K blocks go through one batched sweep (runner.py:326-374 propagates
    # all blocks through each radial window together)
    Ys = propagate_blocks(subs, energies, sectors, propagation) if propagation.mode == "inelastic" and len(subs) > 1 else None
    results = []
    for k, (block, sub) in enumerate(zip(blocks, subs)):
        B, L = get_Bmat_BF_to_SF(sub.basis)
        if Ys is None:
            Y, Smat = solve_device(sub, energies, sectors, propagation, B, L)
        else:
            Y, Smat = match_device(sub, Ys[k], energies, sectors, propagation, B, L)
        indices = np.asarray(block.channel_indices, dtype=np.int64)
        E_int = sub.basis.E_int
        open_idx = tuple(indices[np.flatnonzero(E_int < E)] for E in energies)
        results.append(KBlockResult(block=block, open_channel_indices=open_idx, Y_BF=Y, Bmat=B, L=L,
                                    Smat_asymptotic=Smat))
    timing = Timing(wall_seconds=perf_counter() - wall0, cpu_seconds=process_time() - cpu0)
    logger.info(f"Solver complete: {timing}")
    return CoupledStatesResult(basis=basis, Etot=energies, approx=hamiltonian.approx, blocks=tuple(results), timing=timing)
