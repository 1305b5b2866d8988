"""End-to-end ticc.solve on the GPU against the oracle pipeline (|dS| <= 1e-10)."""

import numpy as np
import pytest

import pyticc_amd as ticc
from oracle import matching as omatch
from oracle import pipeline as opipe
from pyticc_amd import synthetic
from tests.helpers import oracle_problem

pytestmark = pytest.mark.gpu

TOL_S = 1.0e-10


def _compare(workload, n_check=None):
    grid, ham, prob = oracle_problem(workload)
    res = ticc.solve(workload.system, workload.energies, grid, ticc.Propagation(device="gpu"))
    E = workload.energies if n_check is None else workload.energies[:n_check]
    Y, B, L, S = opipe.solve(prob, E)
    np.testing.assert_allclose(res.L, L, rtol=0, atol=1e-12)
    worst = 0.0
    for e in range(E.size):
        assert res.Smat[e].shape == S[e].shape
        if S[e].size:
            worst = max(worst, float(np.max(np.abs(res.Smat[e] - S[e]))))
            assert omatch.unitarity_defect(res.Smat[e]) < 1e-8
            np.testing.assert_allclose(res.Smat[e], res.Smat[e].T, atol=1e-8)
    assert worst <= TOL_S, worst
    return res, worst


def test_solve_atom_rotor_reduced_matches_oracle(gpu_device):
    w = synthetic.atom_rotor(jmax=5, Jtot=4, n_energy=5, boundaries=(3.0, 12.0), half_step=0.02, n_theta=12)
    _compare(w)


def test_solve_rotor_rotor_reduced_matches_oracle(gpu_device):
    w = synthetic.rotor_rotor(jmax_X=2, jmax_Y=3, Jtot=2, n_energy=4, boundaries=(4.0, 10.0), half_step=0.03,
                              n_theta_X=6, n_theta_Y=7, n_phi=6, e_range_cm=(30.37, 300.37))
    _compare(w)


def test_solve_single_channel_zero_potential_is_unitary(gpu_device):
    rovib = ticc.RovibBasis(grids=np.array([1.5]), E_vj=np.zeros((1, 1)), WF_vj=np.ones((1, 1, 1)))
    pes = ticc.PESWrapper(interaction=lambda R, c: np.zeros(c.shape[1]))
    sysm = ticc.build_ScattSystem(ticc.AtomSpec(), ticc.DiatomBasis(rovib=rovib, energy_zero=0.0), Jtot=0, system_parity=1, potential=pes, reduced_mass=2.0)
    grid = ticc.prepare_potential(sysm, (3.0, 3.2), (0.1,), n_theta=4)
    res = ticc.solve(sysm, (0.1, 0.2), grid, ticc.Propagation())
    assert res.basis.n_channel == 1
    assert res.Y_propagated.shape == (2, 1, 1)
    assert res.open_channel_indices[0].tolist() == [0]
    assert res.timing is not None and res.timing.wall_seconds >= 0.0
    np.testing.assert_allclose(np.abs(res.Smat[0]), 1.0, atol=1e-13)
