"""LDMD sweep on the GPU against closed forms and the oracle (reference KATs: tests/unit/test_logd.py)."""

import numpy as np
import pytest
import torch

from oracle import ldmd as oldmd
from pyticc_amd import _device
from pyticc_amd.propagation.logd import initialize_logD_inelastic, propagate_logD, propagate_logD_sector
from tests.helpers import random_W_batch

pytestmark = pytest.mark.gpu


def test_initial_conditions_inelastic(gpu_device):
    Y = initialize_logD_inelastic(np.diag([4.0, -9.0]))
    np.testing.assert_allclose(_device.to_host(Y), np.diag([2.0, 3.0]))
    assert Y.dtype == torch.float64


@pytest.mark.parametrize("W", [4.0, -9.0, 1.0e-12])
def test_sector_is_exact_for_constant_reference_potential(gpu_device, W):
    h, Y0 = 0.15, 1.3
    m = np.array([[W]])
    res = _device.to_host(propagate_logD_sector(np.array([[Y0]]), h, m, m, m))
    np.testing.assert_allclose(res, [[oldmd.constant_logd(Y0, W, 2 * h)]], rtol=1e-13, atol=1e-13)


def test_energy_batch_and_sector_scan_constant_potential(gpu_device):
    E = np.array([0.1, 0.2])
    mu = 2.0
    hs = np.array([0.05, 0.10])
    Wb = np.full((2, 1, 1), 3.0)
    Y0 = np.full((2, 1, 1), 1.25)
    res = _device.to_host(propagate_logD(Y0, E, mu, hs, Wb, Wb, Wb))
    L = 2.0 * hs.sum()
    expect = np.array([[[oldmd.constant_logd(1.25, 3.0 - 2 * mu * e, L)]] for e in E])
    np.testing.assert_allclose(res, expect, rtol=1e-12, atol=1e-12)


def test_sweep_matches_sequential_sector_updates(gpu_device):
    # reference test_logd.py:98-144 with its exact inputs
    E = np.array([0.03, 0.07])
    mu = 1.7
    hs = np.array([0.04, 0.06, 0.05])
    Ws = np.array([[[1.8, 0.10], [0.10, 2.3]], [[1.9, 0.12], [0.12, 2.2]], [[2.0, 0.08], [0.08, 2.1]]])
    Wm = Ws + np.array([[[0.05, 0.02], [0.02, -0.03]]])
    We = Ws + np.array([[[0.09, -0.01], [-0.01, -0.06]]])
    Y0 = np.broadcast_to(np.array([[1.2, 0.03], [0.03, 1.5]]), (2, 2, 2)).copy()
    res = _device.to_host(propagate_logD(Y0, E, mu, hs, Ws, Wm, We))
    seq = []
    for e, y in zip(E, Y0):
        shift = 2 * mu * e * np.eye(2)
        for s in range(3):
            y = _device.to_host(propagate_logD_sector(y, hs[s], Ws[s] - shift, Wm[s] - shift, We[s] - shift))
        seq.append(y)
    np.testing.assert_allclose(res, np.stack(seq), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(res, oldmd.propagate(Y0, E, mu, hs, Ws, Wm, We), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("N,n_E,n_S", [(3, 2, 5), (33, 3, 7), (66, 4, 40), (130, 2, 6), (1024, 2, 2), (1200, 2, 2),
                                        (1536, 2, 2)])
def test_sweep_matches_oracle_random_coupled(gpu_device, N, n_E, n_S):
    rng = np.random.default_rng(N * 7 + n_S)
    Ws, Wm, We = random_W_batch(rng, n_S, N, scale=20.0)
    hs = rng.uniform(0.01, 0.03, n_S)
    E = rng.uniform(0.5, 2.0, n_E)
    mu = 3.0
    W0 = Ws[0][None] - 2 * mu * E[:, None, None] * np.eye(N)
    Y0 = oldmd.initialize_inelastic(W0)
    got = _device.to_host(propagate_logD(Y0, E, mu, hs, Ws, Wm, We))
    ref = oldmd.propagate(Y0, E, mu, hs, Ws, Wm, We)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11 * scale)
    np.testing.assert_array_equal(got, np.swapaxes(got, 1, 2))


def test_bad_half_step_raises(gpu_device):
    with pytest.raises(ValueError, match="positive"):
        propagate_logD(np.eye(2)[None], [0.1], 1.0, [0.0], np.eye(2)[None], np.eye(2)[None], np.eye(2)[None])


def test_nonfinite_result_is_reported(gpu_device):
    W = np.full((1, 1, 1), np.nan)
    with pytest.raises(FloatingPointError):
        propagate_logD(np.ones((1, 1, 1)), [0.1], 1.0, [0.1], W, W, W)


def test_complex_W_is_rejected(gpu_device):
    with pytest.raises(TypeError):
        propagate_logD(np.array([[[1.0j]]]), [0.1], 1.0, [0.1], np.ones((1, 1, 1)) * 1j, np.ones((1, 1, 1)), np.ones((1, 1, 1)))


@pytest.mark.parametrize("groups", ["2", "3"])
def test_energy_groups_give_the_single_group_sweep(gpu_device, monkeypatch, groups):
    """Energies swept in groups on separate streams (TICC_SWEEP_GROUPS) give the
    one-group sweep bit for bit, over 37 sectors (two Q_mid chunks) whose start
    and end points differ between neighbouring sectors."""
    rng = np.random.default_rng(29)
    N, n_E, n_S = 70, 7, 37
    Ws, Wm, We = random_W_batch(rng, n_S, N, scale=20.0)
    hs = rng.uniform(0.01, 0.03, n_S)
    E = rng.uniform(0.5, 2.0, n_E)
    mu = 3.0
    Y0 = oldmd.initialize_inelastic(Ws[0][None] - 2 * mu * E[:, None, None] * np.eye(N))
    one = _device.to_host(propagate_logD(Y0, E, mu, hs, Ws, Wm, We))
    monkeypatch.setenv("TICC_SWEEP_GROUPS", groups)
    grouped = _device.to_host(propagate_logD(Y0, E, mu, hs, Ws, Wm, We))
    np.testing.assert_array_equal(grouped, one)
    ref = oldmd.propagate(Y0, E, mu, hs, Ws, Wm, We)
    np.testing.assert_allclose(grouped, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_zero_sectors_leave_Y_unchanged(gpu_device):
    rng = np.random.default_rng(5)
    N = 9
    Y0 = rng.standard_normal((3, N, N))
    Y0 = Y0 + np.swapaxes(Y0, 1, 2)
    empty = np.zeros((0, N, N))
    got = _device.to_host(propagate_logD(Y0, [0.1, 0.2, 0.3], 2.0, [], empty, empty, empty))
    np.testing.assert_array_equal(got, Y0)


def test_default_two_groups_equal_one_group(gpu_device, monkeypatch):
    """At Np >= 256 with 16 or more energies the sweep runs two energy groups on
    their own streams by default: bit for bit the one-group sweep."""
    rng = np.random.default_rng(31)
    N, n_E, n_S = 300, 16, 5
    Ws, Wm, We = random_W_batch(rng, n_S, N, scale=20.0)
    hs = rng.uniform(0.01, 0.03, n_S)
    E = rng.uniform(0.5, 2.0, n_E)
    mu = 3.0
    Y0 = oldmd.initialize_inelastic(Ws[0][None] - 2 * mu * E[:, None, None] * np.eye(N))
    print(f"default groups: N={N} energies={n_E} sectors={n_S}", flush=True)
    default = _device.to_host(propagate_logD(Y0, E, mu, hs, Ws, Wm, We))
    monkeypatch.setenv("TICC_SWEEP_GROUPS", "1")
    one = _device.to_host(propagate_logD(Y0, E, mu, hs, Ws, Wm, We))
    np.testing.assert_array_equal(default, one)
