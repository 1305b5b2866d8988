"""Capture mode on the device (Propagation(mode="capture"); src/pyticc/propagation/logd.py:72-91,
runner.py:284-291, smatrix.py:139-223 with complex Y): the incoming-wave boundary
condition, the reference's capture KATs (test_logd.py:58-67 incoming wave
invariant; test_smatrix.py:148-161 capture S = 2/3), a coupled complex sweep and
a whole capture solve against the oracle."""

import numpy as np
import pytest

import pyticc_amd as ticc
from oracle import bessel as ob
from oracle import ldmd, pipeline
from oracle import matching as om
from pyticc_amd import synthetic
from pyticc_amd.match.smatrix import get_Smat
from pyticc_amd.propagation.logd import initialize_logD_capture, propagate_logD
from tests.helpers import oracle_problem

pytestmark = pytest.mark.gpu

TOL_S = 1.0e-10


def _host(t):
    return t.detach().cpu().numpy()


def test_capture_initial_condition(gpu_device):
    W = np.array([[4.0, 0.3, 0.0], [0.3, -9.0, 0.1], [0.0, 0.1, 0.0]])
    got = _host(initialize_logD_capture(W))
    np.testing.assert_allclose(got, ldmd.initialize_capture(W), rtol=0, atol=1e-15)
    np.testing.assert_allclose(got, np.diag([2.0, -3.0j, 0.0]), rtol=0, atol=1e-15)


def test_capture_incoming_wave_invariant(gpu_device):
    W = np.array([[[-9.0]]])
    got = _host(propagate_logD(np.array([[[-3.0j]]]), [0.0], 0.5, [0.1], W, W, W))
    np.testing.assert_allclose(got, [[[-3.0j]]], rtol=1e-13, atol=1e-13)


def test_complex_sweep_matches_oracle(gpu_device):
    rng = np.random.default_rng(11)
    n, nE, nS = 5, 3, 7
    base = rng.standard_normal((n, n))
    base = 0.5 * (base + base.T) + np.diag([40.0, 10.0, -5.0, -20.0, -60.0])
    Ws, Wm, We = (base[None] + 0.05 * k * rng.standard_normal((nS, n, n)) for k in (1, 2, 3))
    Ws, Wm, We = (0.5 * (W + np.swapaxes(W, 1, 2)) for W in (Ws, Wm, We))
    E = np.array([0.0, 0.3, 1.1])
    mu = 2.5
    hs = np.full(nS, 0.04)
    Y0 = np.stack([ldmd.initialize_capture(Ws[0] - 2 * mu * e * np.eye(n)) for e in E])
    got = _host(propagate_logD(Y0, E, mu, hs, Ws, Wm, We))
    want = Y0.copy()
    for s in range(nS):
        ref, Qs, Qm, Qe = ldmd.correction_matrices(hs[s], Ws[s], Wm[s], We[s])
        for e in range(nE):
            want[e] = ldmd.sector_step(want[e], hs[s], ref - 2 * mu * E[e], Qs, Qm, Qe)
    assert np.abs(want.imag).max() > 0.1
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())


def test_capture_S_two_thirds(gpu_device):
    """Single open channel, K = -0.2 i: S = (1 + iK)^-1 (1 - iK) = 0.8 / 1.2."""
    k = np.sqrt(2 * 2.0 * 0.4)
    j, nn, jd, nd = ob.riccati_bessel_jy(0.0, k * 6.0)
    J, N, Jp, Np = j / np.sqrt(k), nn / np.sqrt(k), np.sqrt(k) * jd, np.sqrt(k) * nd
    K = -0.2j
    Y = np.array([[[(Jp + Np * K) / (J + N * K)]]])
    (S,) = get_Smat(Y, 6.0, [0.4], 2.0, [0.0], [0.0])
    np.testing.assert_allclose(S, [[2.0 / 3.0]], atol=1e-12)


def test_capture_solve_matches_oracle(gpu_device):
    w = synthetic.atom_rotor(jmax=4, Jtot=2, n_energy=3, boundaries=(5.0, 12.0), half_step=0.02)
    grid, ham, prob = oracle_problem(w)
    res = ticc.solve(w.system, w.energies, grid, ticc.Propagation(mode="capture", device="gpu"))
    Y, B, L, S = pipeline.solve(prob, w.energies, mode="capture")
    assert np.iscomplexobj(res.Y_propagated)
    np.testing.assert_allclose(res.Y_propagated, Y, rtol=0, atol=1e-10 * np.abs(Y).max())
    worst = 0.0
    for got, want in zip(res.Smat, S):
        assert got.shape == want.shape
        worst = max(worst, float(np.max(np.abs(got - want))))
    assert worst <= TOL_S, worst
    defects = [om.unitarity_defect(s) for s in res.Smat]
    assert max(defects) > 0.1  # flux is absorbed at the inner boundary
    np.testing.assert_allclose(defects, [om.unitarity_defect(s) for s in S], atol=1e-9)
