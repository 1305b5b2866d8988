"""Shared builders: product setup -> oracle Problem, random coupled W batches."""

from __future__ import annotations

import numpy as np

from oracle.pipeline import Problem
from tests.golden.fixtures import (  # noqa: F401
    n800_match_tolerance, n800_tolerance, n800_total_tolerance, rotor_rotor_N800_reference)
from pyticc_amd import prepare_potential
from pyticc_amd.matrix.centrifugal import get_Umat_BF, ladder_labels
from pyticc_amd.scattering.hamiltonian import build_hamiltonian


def oracle_problem(workload):
    """Prepare the PES grid and the oracle Problem for a synthetic workload.  The
    Problem takes the product's setup arrays (channels, U, B_K rows, PES grid); for
    the benchmark systems tests/test_oracle_setup.py shows them equal to the
    oracle's own restatement (oracle/setup.py)."""
    grid = prepare_potential(workload.system, workload.boundaries, workload.half_steps, **workload.quadrature)
    ham = build_hamiltonian(workload.system, grid)
    return grid, ham, hamiltonian_problem(ham, grid)


def hamiltonian_problem(ham, grid):
    """Oracle Problem of a (possibly block-restricted) Hamiltonian on a prepared grid."""
    vb = ham.vbasis
    blocks = [(np.asarray(vb.channel_indices[K]), vb.B_real[K], None if vb.B_imag is None else vb.B_imag[K]) for K in sorted(vb.channel_indices)]
    basis = ham.basis
    return Problem(
        blocks=blocks,
        norm=vb.normalization,
        U=get_Umat_BF(basis),
        E_int=basis.E_int,
        mu=ham.reduced_mass,
        ladder=ladder_labels(basis),
        j_couple=np.asarray([c.j_couple for c in basis]),
        Jtot=basis.Jtot,
        sectors=[(s.radial_start, s.radial_end) for s in grid.sectors],
        V=grid.values.reshape(grid.radial_points.size, -1),
    )


def random_W_batch(rng, n_sector, N, scale=1.0, diag_shift=0.0):
    """Symmetric energy-independent W_base batches (start, mid, end) that vary smoothly."""
    base = rng.standard_normal((N, N)) * 0.3 * scale
    base = 0.5 * (base + base.T) + np.diag(rng.uniform(-2.0, 4.0, N) * scale + diag_shift)
    out = []
    for _ in range(3):
        d = rng.standard_normal((n_sector, N, N)) * 0.05 * scale
        out.append(base[None] + 0.5 * (d + np.swapaxes(d, 1, 2)))
    return out

