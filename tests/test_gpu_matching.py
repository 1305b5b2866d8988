"""Device K->S matching against the reference KATs (tests/unit/test_smatrix.py) and the oracle."""

import numpy as np
import pytest

from oracle import bessel as obessel
from oracle import matching as omatch
from pyticc_amd import _device
from pyticc_amd.match.smatrix import get_Smat, get_Smat_device, unpack_smat
from pyticc_amd.propagation.logd import propagate_logD

pytestmark = pytest.mark.gpu


def _ref_mats(energy, Rmatch, mu, E_int, L):
    J = np.ones(E_int.size)
    N = np.ones(E_int.size)
    Jp = np.empty(E_int.size)
    Np = np.empty(E_int.size)
    for i, thr in enumerate(E_int):
        k = np.sqrt(2 * mu * abs(energy - thr))
        if energy > thr:
            j, n, jd, nd = obessel.riccati_bessel_jy(float(L[i]), k * Rmatch)
            J[i], N[i], Jp[i], Np[i] = j / np.sqrt(k), n / np.sqrt(k), np.sqrt(k) * jd, np.sqrt(k) * nd
        else:
            Jp[i] = 0.5 / Rmatch + k * obessel.modified_bessel_I_logD(float(L[i] + 0.5), k * Rmatch)
            Np[i] = 0.5 / Rmatch + k * obessel.modified_bessel_K_logD(float(L[i] + 0.5), k * Rmatch)
    return np.diag(J), np.diag(N), np.diag(Jp), np.diag(Np)


def _Y_from_K(J, N, Jp, Np, K):
    return (Jp + Np @ K) @ np.linalg.inv(J + N @ K)


def test_free_single_channel_is_identity(gpu_device):
    J, N, Jp, Np = _ref_mats(0.4, 6.0, 2.0, np.array([0.0]), np.array([0.0]))
    Y = _Y_from_K(J, N, Jp, Np, np.zeros((1, 1)))
    (S,) = get_Smat(Y[None], 6.0, [0.4], 2.0, [0.0], [0.0])
    np.testing.assert_allclose(S, np.eye(1), rtol=1e-13, atol=1e-13)


def test_free_propagation_and_matching_is_identity(gpu_device):
    E, mu = 0.2, 2.0
    k = np.sqrt(2 * mu * E)
    W = np.zeros((5, 1, 1))
    Y0 = np.array([[[k / np.tan(k * 1.1)]]])
    Y = propagate_logD(Y0, [E], mu, np.full(5, 0.1), W, W, W)
    (S,) = get_Smat(Y, 2.1, [E], mu, [0.0], [0.0])
    np.testing.assert_allclose(S, np.eye(1), rtol=1e-12, atol=1e-12)


def test_real_open_reaction_matrix_closed_form_and_unitary(gpu_device):
    E_int, L = np.array([0.0, 0.1]), np.array([0.0, 1.3])
    K = np.array([[0.2, 0.07], [0.07, -0.1]])
    J, N, Jp, Np = _ref_mats(0.5, 7.0, 2.0, E_int, L)
    (S,) = get_Smat(_Y_from_K(J, N, Jp, Np, K)[None], 7.0, [0.5], 2.0, E_int, L)
    I = np.eye(2)
    np.testing.assert_allclose(S, np.linalg.solve(I + 1j * K, I - 1j * K), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(S.conj().T @ S, I, rtol=1e-12, atol=1e-12)


def test_closed_channel_eliminated_before_open_block(gpu_device):
    E_int, L = np.array([0.0, 0.5]), np.array([0.0, 1.2])
    K = np.array([[0.25, 0.08], [0.08, -0.3]])
    J, N, Jp, Np = _ref_mats(0.2, 8.0, 3.0, E_int, L)
    (S,) = get_Smat(_Y_from_K(J, N, Jp, Np, K)[None], 8.0, [0.2], 3.0, E_int, L)
    assert S.shape == (1, 1)
    np.testing.assert_allclose(S[0, 0], (1 - 0.25j) / (1 + 0.25j), rtol=1e-12, atol=1e-12)


def test_elimination_matches_full_reaction_matrix(gpu_device):
    E_int, L = np.array([0.0, 0.12, 0.55, 1.1]), np.array([0.0, 1.0, 2.3, 4.2])
    Y = np.array([[0.4, 0.03, 0.08, -0.02], [0.03, -0.2, 0.04, 0.06], [0.08, 0.04, -1.1, 0.05], [-0.02, 0.06, 0.05, -1.4]])
    J, N, Jp, Np = _ref_mats(0.3, 9.0, 2.5, E_int, L)
    Kf = -np.linalg.solve(Y @ N - Np, Y @ J - Jp)
    I = np.eye(2)
    expect = np.linalg.solve(I + 1j * Kf[:2, :2], I - 1j * Kf[:2, :2])
    (S,) = get_Smat(Y[None], 9.0, [0.3], 2.5, E_int, L)
    np.testing.assert_allclose(S, expect, rtol=2e-13, atol=2e-13)


def test_energy_at_threshold_raises(gpu_device):
    with pytest.raises(ValueError, match="threshold"):
        get_Smat(np.eye(2)[None], 5.0, [0.1], 1.0, [0.1, 0.0], [0.0, 1.0])


@pytest.mark.parametrize("N,E,tol", [(70, (0.05, 0.5, 1.0, 1.5, 2.5, 0.77), 1e-11), (1100, (0.5, 1.5), 1e-10)])
def test_matching_matches_oracle_random_batch(gpu_device, N, E, tol):
    rng = np.random.default_rng(11)
    E = np.asarray(E)
    A = rng.standard_normal((E.size, N, N))
    Y = 0.5 * (A + np.swapaxes(A, 1, 2)) + np.diag(rng.uniform(1.0, 5.0, N))
    E_int = np.sort(rng.uniform(0.0, 2.0, N))
    L = rng.integers(0, 12, N) + rng.choice([0.0, 0.3], N)
    B = np.linalg.qr(rng.standard_normal((N, N)))[0]
    S_re, S_im, n_open = get_Smat_device(Y, B, 12.0, E, 3.0, E_int, L, gpu_device)
    got = unpack_smat(S_re, S_im, n_open)
    ref = omatch.get_smat(omatch.transform_bf_to_sf(Y, B), 12.0, E, 3.0, E_int, L)
    for g, r in zip(got, ref):
        assert g.shape == r.shape
        np.testing.assert_allclose(g, r, rtol=0, atol=tol)


def test_ragged_batch_with_every_channel_closed(gpu_device):
    """An energy below every threshold gives an empty S (0, 0) beside a (2, 2) one,
    as the oracle's open/closed split (smatrix.py:56-57) does."""
    rng = np.random.default_rng(1)
    N = 5
    A = rng.standard_normal((2, N, N))
    Y = 0.5 * (A + np.swapaxes(A, 1, 2)) + np.diag(rng.uniform(1.0, 5.0, N))
    E_int = np.array([0.3, 0.5, 0.7, 0.9, 1.1])
    L = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    E = np.array([0.1, 0.6])
    S_re, S_im, n_open = get_Smat_device(Y, np.eye(N), 12.0, E, 3.0, E_int, L, gpu_device)
    got = unpack_smat(S_re, S_im, n_open)
    ref = omatch.get_smat(Y, 12.0, E, 3.0, E_int, L)
    assert [g.shape for g in got] == [(0, 0), (2, 2)]
    for g, r in zip(got, ref):
        assert g.shape == r.shape
        np.testing.assert_allclose(g, r, rtol=0, atol=1e-12)
