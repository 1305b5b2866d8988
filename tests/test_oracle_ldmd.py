"""Pins of the oracle LDMD propagator and sector layout (CPU).

Closed forms and invariants, not a retyping of the oracle's formula:
constant single-channel W is exact (tanh/tan), the capture incoming wave is
invariant, a coupled constant W converges to the exact eigenbasis propagator
at O(h^4), batched sweep == sequential sectors, sector layout KATs.
"""

import numpy as np
import pytest

from oracle import grid as ogrid
from oracle import ldmd


def test_initial_conditions():
    W = np.diag([4.0, -9.0])
    np.testing.assert_allclose(ldmd.initialize_inelastic(W), np.diag([2.0, 3.0]))
    np.testing.assert_allclose(ldmd.initialize_capture(W), np.diag([2.0, -3.0j]))


@pytest.mark.parametrize("W", [4.0, -9.0, 0.0, 1e-13, -2.5e3])
def test_constant_single_channel_is_exact(W):
    h, Y0 = 0.15, 1.3
    ref, Qs, Qm, Qe = ldmd.correction_matrices(h, np.array([[W]]), np.array([[W]]), np.array([[W]]))
    got = ldmd.sector_step(np.array([[Y0]]), h, ref, Qs, Qm, Qe)
    expect = ldmd.constant_logd(Y0, W, 2 * h) if W != 0.0 else Y0 / (1 + 2 * h * Y0)
    np.testing.assert_allclose(got, [[expect]], rtol=1e-13, atol=1e-13)


def test_capture_incoming_wave_invariant():
    W = np.array([[-9.0]])
    ref, Qs, Qm, Qe = ldmd.correction_matrices(0.1, W, W, W)
    got = ldmd.sector_step(np.array([[-3.0j]]), 0.1, ref, Qs, Qm, Qe)
    np.testing.assert_allclose(got, [[-3.0j]], rtol=1e-13, atol=1e-13)


def _exact_constant_coupled(Y0, W, L):
    # eigenbasis of a constant W: each eigenchannel propagates with the closed form
    w, O = np.linalg.eigh(W)
    Yt = O.T @ Y0 @ O
    # propagate the full matrix Riccati equation Y' = W - Y^2 exactly via the
    # transfer matrix of psi'' = W psi: [psi, psi'](L) = M [psi, psi'](0)
    c = np.where(w > 0, np.cosh(np.sqrt(np.abs(w)) * L), np.cos(np.sqrt(np.abs(w)) * L))
    s = np.where(w > 0, np.sinh(np.sqrt(np.abs(w)) * L) / np.sqrt(np.abs(w)), np.sin(np.sqrt(np.abs(w)) * L) / np.sqrt(np.abs(w)))
    sd = np.where(w > 0, np.sqrt(np.abs(w)) * np.sinh(np.sqrt(np.abs(w)) * L), -np.sqrt(np.abs(w)) * np.sin(np.sqrt(np.abs(w)) * L))
    psi = np.diag(c) + np.diag(s) @ Yt
    dpsi = np.diag(sd) + np.diag(c) @ Yt
    Yl = dpsi @ np.linalg.inv(psi)
    return O @ Yl @ O.T


def test_coupled_constant_W_converges_at_fourth_order():
    rng = np.random.default_rng(3)
    N = 4
    A = rng.standard_normal((N, N))
    W = 0.5 * (A + A.T) + np.diag([3.0, -2.0, 5.0, -1.0])
    Y0 = np.diag([1.0, 0.5, 2.0, 0.8]) + 0.05 * (A + A.T)
    L = 0.6
    exact = _exact_constant_coupled(Y0, W, L)
    errs = []
    for n in (8, 16, 32):
        h = L / (2 * n)
        Y = Y0.copy()
        for _ in range(n):
            ref, Qs, Qm, Qe = ldmd.correction_matrices(h, W, W, W)
            Y = ldmd.sector_step(Y, h, ref, Qs, Qm, Qe)
        errs.append(np.max(np.abs(Y - exact)))
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(orders > 3.6), (errs, orders)
    assert errs[-1] < 1e-6


def test_sweep_equals_sequential_sectors():
    rng = np.random.default_rng(4)
    N, nS = 3, 4
    E = np.array([0.03, 0.07])
    mu = 1.7
    hs = rng.uniform(0.03, 0.06, nS)
    base = rng.standard_normal((N, N))
    Ws = np.stack([0.5 * (base + base.T) + np.diag([2.0, 3.0, 1.0]) + 0.01 * s for s in range(nS)])
    Wm, We = Ws + 0.02, Ws + 0.04
    Y0 = np.broadcast_to(np.diag([1.2, 1.5, 0.9]), (2, N, N)).copy()
    got = ldmd.propagate(Y0, E, mu, hs, Ws, Wm, We)
    for e in range(2):
        Y = Y0[e]
        for s in range(nS):
            shift = 2 * mu * E[e] * np.eye(N)
            ref, Qs, Qm, Qe = ldmd.correction_matrices(hs[s], Ws[s] - shift, Wm[s] - shift, We[s] - shift)
            Y = ldmd.sector_step(Y, hs[s], ref, Qs, Qm, Qe)
        np.testing.assert_allclose(got[e], Y, rtol=1e-12, atol=1e-12)


def test_sector_layout_kats():
    s = ogrid.build_radial_sectors([3.0, 3.5, 4.1], [0.1, 0.2])
    np.testing.assert_allclose([a for a, _ in s], [3.0, 3.2, 3.4, 3.5, 3.9])
    np.testing.assert_allclose([c for _, c in s], [3.2, 3.4, 3.5, 3.9, 4.1])
    np.testing.assert_allclose(ogrid.half_steps_of(s), [0.1, 0.1, 0.05, 0.2, 0.1])
    s = ogrid.build_radial_sectors([0.0, 1.0], [0.05])
    assert len(s) == 10 and s[-1][1] == 1.0
    assert len(ogrid.build_radial_sectors([4.0, 40.0], [0.009])) == 2000
    assert len(ogrid.build_radial_sectors([3.0, 30.0], [0.01])) == 1350
    pts = ogrid.radial_points(ogrid.build_radial_sectors([3.0, 3.2], [0.1]))
    np.testing.assert_allclose(pts, [3.0, 3.1, 3.2])


@pytest.mark.parametrize("mistake", ["qmid_sign", "drop_qstart", "qmid_factor"])
def test_fourth_order_check_catches_planted_mistakes(monkeypatch, mistake):
    """The convergence pin fails for plausible mistakes planted in the oracle."""
    good = ldmd.correction_matrices

    def planted(h, Ws, Wm, We):
        ref, Qs, Qm, Qe = good(h, Ws, Wm, We)
        if mistake == "qmid_sign":
            Qm = -Qm
        elif mistake == "drop_qstart":
            Qs = 0.0 * Qs
        else:
            Qm = 0.5 * Qm
        return ref, Qs, Qm, Qe

    monkeypatch.setattr(ldmd, "correction_matrices", planted)
    with pytest.raises(AssertionError):
        test_coupled_constant_W_converges_at_fourth_order()


@pytest.mark.parametrize("n", [1, 7, 32, 33, 96, 161])
def test_gauss_jordan_inverse_matches_lapack(n):
    """The oracle's blocked Gauss-Jordan inverse (the device's algorithm class)
    against LAPACK's inverse, on random matrices, a permutation matrix and a
    matrix with a zero diagonal (every column needs a row interchange)."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 0.1 * np.eye(n)
    ref = np.linalg.inv(A)
    np.testing.assert_allclose(ldmd.gauss_jordan_inverse(A), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    Pm = np.eye(n)[rng.permutation(n)]
    np.testing.assert_array_equal(ldmd.gauss_jordan_inverse(Pm), Pm.T)
    Z = rng.standard_normal((n, n))
    np.fill_diagonal(Z, 0.0)
    if n > 1:
        np.testing.assert_allclose(ldmd.gauss_jordan_inverse(Z) @ Z, np.eye(n), rtol=0, atol=1e-11)


def test_gauss_jordan_inverse_block_size_independent():
    """Blocking changes only the rounding: nb = 1 (column by column), 8, 32 and a
    block wider than the matrix give the same inverse to rounding."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((70, 70))
    ref = ldmd.gauss_jordan_inverse(A, nb=1)
    for nb in (8, 32, 128):
        np.testing.assert_allclose(ldmd.gauss_jordan_inverse(A, nb=nb), ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("n", [2, 64, 256])
def test_gauss_jordan_inverse_hadamard_exact(n):
    """Sylvester Hadamard matrices (every pivot column tied in magnitude): the
    Gauss-Jordan inverse is exactly H^T / n at every block size."""
    H = np.array([[1.0]])
    while H.shape[0] < n:
        H = np.block([[H, H], [H, -H]])
    for nb in (1, 32):
        np.testing.assert_array_equal(ldmd.gauss_jordan_inverse(H, nb=nb), H.T / n)


def test_sector_orderings_agree():
    """The solve-based, explicit-inverse and Gauss-Jordan sector steps are FP64
    orderings of the same arithmetic: one coupled sector agrees to rounding."""
    rng = np.random.default_rng(11)
    n, h = 24, 0.02
    B = rng.standard_normal((n, n)) * 0.3
    Wm = 0.5 * (B + B.T) + np.diag(rng.uniform(-2.0, 4.0, n))
    Ws, We = Wm + 0.01, Wm - 0.01
    ref, Qs, Qm, Qe = ldmd.correction_matrices(h, Ws, Wm, We)
    Y = np.diag(rng.uniform(0.5, 2.0, n))
    a = ldmd.sector_step(Y, h, ref, Qs, Qm, Qe)
    for step in (ldmd.sector_step_inverse, ldmd.sector_step_gj):
        np.testing.assert_allclose(step(Y, h, ref, Qs, Qm, Qe), a, rtol=0, atol=1e-11 * np.abs(a).max())
