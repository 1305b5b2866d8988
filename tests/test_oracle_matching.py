"""Pins of the oracle matching (CPU): reference KATs of tests/unit/test_smatrix.py,
test_bessel.py, test_asymptotic.py, plus closed forms for the cross section."""

import numpy as np
import pytest
from scipy import special

from oracle import bessel as ob
from oracle import ldmd
from oracle import matching as om


def _ref_mats(energy, Rmatch, mu, E_int, L):
    J = np.ones(E_int.size)
    N = np.ones(E_int.size)
    Jp = np.empty(E_int.size)
    Np = np.empty(E_int.size)
    for i, thr in enumerate(E_int):
        k = np.sqrt(2 * mu * abs(energy - thr))
        if energy > thr:
            j, n, jd, nd = ob.riccati_bessel_jy(float(L[i]), k * Rmatch)
            J[i], N[i], Jp[i], Np[i] = j / np.sqrt(k), n / np.sqrt(k), np.sqrt(k) * jd, np.sqrt(k) * nd
        else:
            Jp[i] = 0.5 / Rmatch + k * ob.modified_bessel_I_logD(float(L[i] + 0.5), k * Rmatch)
            Np[i] = 0.5 / Rmatch + k * ob.modified_bessel_K_logD(float(L[i] + 0.5), k * Rmatch)
    return np.diag(J), np.diag(N), np.diag(Jp), np.diag(Np)


def _Y(J, N, Jp, Np, K):
    return (Jp + Np @ K) @ np.linalg.inv(J + N @ K)


def test_bessel_elementary_wronskian_and_scipy():
    j, n, jd, nd = ob.riccati_bessel_jy(0.0, 1.7)
    np.testing.assert_allclose([j, n, jd, nd], [np.sin(1.7), -np.cos(1.7), np.cos(1.7), np.sin(1.7)], rtol=1e-14, atol=1e-14)
    j, n, jd, nd = ob.riccati_bessel_jy(2.3, 4.2)
    np.testing.assert_allclose(j * nd - jd * n, 1.0, rtol=1e-13)
    np.testing.assert_allclose(ob.modified_bessel_K_logD(1.7, 3.1), special.kvp(1.7, 3.1) / special.kv(1.7, 3.1), rtol=1e-13)
    assert np.isfinite(ob.modified_bessel_K_logD(48.5, 2500.0))
    # spherical closed form: j_1 riccati = sin x / x - cos x
    x = 3.3
    np.testing.assert_allclose(ob.riccati_bessel_jy(1.0, x)[0], np.sin(x) / x - np.cos(x), rtol=1e-14)


def test_free_channel_is_identity():
    J, N, Jp, Np = _ref_mats(0.4, 6.0, 2.0, np.array([0.0]), np.array([0.0]))
    (S,) = om.get_smat(_Y(J, N, Jp, Np, np.zeros((1, 1)))[None], 6.0, [0.4], 2.0, [0.0], [0.0])
    np.testing.assert_allclose(S, np.eye(1), atol=1e-13)


def test_free_propagation_then_matching_is_identity():
    E, mu = 0.2, 2.0
    k = np.sqrt(2 * mu * E)
    W = np.zeros((5, 1, 1))
    Y = ldmd.propagate(np.array([[[k / np.tan(k * 1.1)]]]), [E], mu, np.full(5, 0.1), W, W, W)
    (S,) = om.get_smat(Y, 2.1, [E], mu, [0.0], [0.0])
    np.testing.assert_allclose(S, np.eye(1), atol=1e-12)


def test_k_to_s_closed_form_and_closed_elimination():
    E_int, L = np.array([0.0, 0.1]), np.array([0.0, 1.3])
    K = np.array([[0.2, 0.07], [0.07, -0.1]])
    J, N, Jp, Np = _ref_mats(0.5, 7.0, 2.0, E_int, L)
    (S,) = om.get_smat(_Y(J, N, Jp, Np, K)[None], 7.0, [0.5], 2.0, E_int, L)
    I = np.eye(2)
    np.testing.assert_allclose(S, np.linalg.solve(I + 1j * K, I - 1j * K), atol=1e-12)
    assert om.unitarity_defect(S) < 1e-12
    E_int, L = np.array([0.0, 0.5]), np.array([0.0, 1.2])
    K = np.array([[0.25, 0.08], [0.08, -0.3]])
    J, N, Jp, Np = _ref_mats(0.2, 8.0, 3.0, E_int, L)
    (S,) = om.get_smat(_Y(J, N, Jp, Np, K)[None], 8.0, [0.2], 3.0, E_int, L)
    np.testing.assert_allclose(S[0, 0], (1 - 0.25j) / (1 + 0.25j), atol=1e-12)


def test_capture_boundary_nonunitary():
    J, N, Jp, Np = _ref_mats(0.4, 6.0, 2.0, np.array([0.0]), np.array([0.0]))
    (S,) = om.get_smat(_Y(J, N, Jp, Np, np.array([[-0.2j]]))[None], 6.0, [0.4], 2.0, [0.0], [0.0])
    np.testing.assert_allclose(S[0, 0], 2.0 / 3.0, atol=1e-12)


def test_threshold_energy_raises():
    with pytest.raises(ValueError, match="threshold"):
        om.reference_functions(0.1, 5.0, 1.0, np.array([0.1]), np.array([0.0]))


def test_bmat_diagonalises_ladders():
    # atom + j=1 rotor, J=2: K in {0,1} ladder, U from the centrifugal closed form
    J, j = 2, 1
    lp = lambda a, b: np.sqrt(a * (a + 1) - b * (b + 1))  # noqa: E731
    U = np.array([[J * (J + 1) + j * (j + 1), -np.sqrt(2) * lp(J, 0) * lp(j, 0)], [-np.sqrt(2) * lp(J, 0) * lp(j, 0), J * (J + 1) + j * (j + 1) - 2]])
    B, L = om.bmat_bf_to_sf(U, np.array([0, 0]), np.array([1, 1]), J)
    np.testing.assert_allclose(B.T @ U @ B, np.diag(L * (L + 1)), atol=1e-13)
    np.testing.assert_allclose(L, [1.0, 3.0], atol=1e-13)
    assert np.all(B[-1] >= 0.0)


def test_partial_cross_section_single_channel_phase_shift():
    delta, k = 0.37, 1.3
    S = np.array([[np.exp(2j * delta)]])
    _, sig = om.partial_cross_sections(S, np.array([0]), np.array([k]), {0: 1}, 0)
    np.testing.assert_allclose(sig[0, 0], 4 * np.pi / k**2 * np.sin(delta) ** 2, rtol=1e-14)


def test_partial_cross_section_detailed_balance():
    rng = np.random.default_rng(9)
    K = rng.standard_normal((5, 5))
    K = K + K.T
    I = np.eye(5)
    S = np.linalg.solve(I + 1j * K, I - 1j * K)  # symmetric unitary
    levels = np.array([0, 0, 1, 1, 1])
    k = np.array([1.1, 1.1, 0.7, 0.7, 0.7])
    g = {0: 2, 1: 3}
    labs, sig = om.partial_cross_sections(S, levels, k, g, 3)
    kk = np.array([1.1, 0.7])
    gg = np.array([2, 3])
    np.testing.assert_allclose(kk[:, None] ** 2 * gg[:, None] * sig, (kk[:, None] ** 2 * gg[:, None] * sig).T, rtol=1e-12)


def test_matching_orderings_agree():
    """The matching with its solves formed through explicit inverses (the ordering
    whose distance sets the matching's noise floor) equals the solve-based S to
    rounding, with closed channels eliminated."""
    rng = np.random.default_rng(21)
    N = 30
    A = rng.standard_normal((2, N, N))
    Y = 0.5 * (A + np.swapaxes(A, 1, 2)) + np.diag(rng.uniform(1.0, 5.0, N))
    E_int = np.sort(rng.uniform(0.0, 2.0, N))
    L = rng.integers(0, 8, N).astype(float)
    E = np.array([0.7, 1.6])
    a = om.get_smat(Y, 12.0, E, 3.0, E_int, L)
    b = om.get_smat(Y, 12.0, E, 3.0, E_int, L, inverse=True)
    for x, y in zip(a, b):
        assert x.shape == y.shape and x.size
        np.testing.assert_allclose(y, x, rtol=0, atol=1e-12)
