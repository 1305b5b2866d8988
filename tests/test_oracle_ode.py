"""The oracle's LDMD propagation against an independent integration of the
coupled equations psi'' = (W_base - 2 mu E) psi (CPU).

The reference holds no S-matrix of a nontrivial system, so its propagator
(propagation/logd.py:143-229, driven by runner.py:165-322) is pinned here by the
mathematics it discretises: for an R-dependent coupled W the sector step must
converge at O(h^4) to the solution of the linear ODE, here integrated by
scipy's DOP853 and by classical RK4 (a different method on the same W points),
and the composed S-matrix of a real multichannel system must converge to the
RK4 limit.  The same checks fail for plausible mistakes planted in the oracle:
start and end corrections swapped, the sign of the midpoint correction flipped,
the reference potential taken at the sector start.
"""

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle import grid as ogrid
from oracle import ldmd
from oracle import matching as omatch
from oracle import pipeline as opipe
from pyticc_amd import synthetic
from tests.helpers import oracle_problem

_L = np.array([0, 1, 2, 3])
_THR = np.array([-6.0, -2.5, 1.0, 4.0])  # two open, two closed channels
_C = np.array([[1.0, 0.4, -0.3, 0.2], [0.4, -0.5, 0.6, -0.1], [-0.3, 0.6, 0.8, 0.5], [0.2, -0.1, 0.5, -0.7]])


def _W(R):
    return np.diag(_THR + _L * (_L + 1) / R**2) + 4.0 * _C * np.exp(-0.7 * (R - 3.0)) * np.cos(1.3 * R)


def _ode_log_derivative(Y0, R0, R1):
    """Y(R1) = psi' psi^-1 with psi(R0) = I, psi'(R0) = Y0, by DOP853 at tolerance 1e-13."""
    n = Y0.shape[0]

    def rhs(R, y):
        psi = y[: n * n].reshape(n, n)
        return np.concatenate([y[n * n:], (_W(R) @ psi).ravel()])

    y0 = np.concatenate([np.eye(n).ravel(), Y0.ravel()])
    sol = solve_ivp(rhs, (R0, R1), y0, method="DOP853", rtol=1e-13, atol=1e-13)
    assert sol.success
    psi = sol.y[: n * n, -1].reshape(n, n)
    dpsi = sol.y[n * n:, -1].reshape(n, n)
    return dpsi @ np.linalg.inv(psi)


_CORRECTIONS = ldmd.correction_matrices


def _variant_corrections(variant):
    def corrections(h, Ws, Wm, We):
        if variant == "ref_at_start":
            return _CORRECTIONS(h, Ws, Ws, We)
        ref, Qs, Qm, Qe = _CORRECTIONS(h, Ws, Wm, We)
        if variant == "swap_start_end":
            Qs, Qe = Qe, Qs
        elif variant == "flip_mid_sign":
            Qm = -Qm
        return ref, Qs, Qm, Qe

    return corrections


def _ldmd_log_derivative(Y0, R0, R1, n_sector, corrections=ldmd.correction_matrices):
    Y = Y0.copy()
    h = (R1 - R0) / (2 * n_sector)
    for i in range(n_sector):
        Rs = R0 + 2 * h * i
        Y = ldmd.sector_step(Y, h, *corrections(h, _W(Rs), _W(Rs + h), _W(Rs + 2 * h)))
    return Y


def _orders(errs):
    return np.log2(np.asarray(errs[:-1]) / np.asarray(errs[1:]))


_Y0 = np.diag([2.0, 1.5, 1.0, 2.5]) + 0.1


def test_sector_step_converges_to_the_coupled_ode_at_fourth_order():
    exact = _ode_log_derivative(_Y0, 3.0, 5.0)
    errs = [np.abs(_ldmd_log_derivative(_Y0, 3.0, 5.0, n) - exact).max() for n in (20, 40, 80, 160)]
    assert np.all(_orders(errs) > 3.8), errs
    assert errs[-1] < 5e-9 * np.abs(exact).max()


@pytest.mark.parametrize("variant", ["swap_start_end", "flip_mid_sign", "ref_at_start"])
def test_planted_mistakes_fail_the_ode_check(variant):
    exact = _ode_log_derivative(_Y0, 3.0, 5.0)
    wrong = _variant_corrections(variant)
    errs = [np.abs(_ldmd_log_derivative(_Y0, 3.0, 5.0, n, wrong) - exact).max() for n in (20, 40, 80)]
    assert np.all(_orders(errs) < 2.5) and errs[-1] > 1e-5, errs


def _rk4_log_derivative(p, energies):
    """Y_BF at the last sector end by classical RK4 on psi'' = (W - 2 mu E) psi, one
    step per sector (W at its start, midpoint and end), with the columns of
    [psi; psi'] re-orthonormalised after every step (Y = psi' psi^-1 is unchanged
    by any right factor)."""
    hs = ogrid.half_steps_of(p.sectors)
    W = opipe.w_points(p, np.arange(2 * len(p.sectors) + 1))
    N = p.n_channel
    eye = np.eye(N)
    shift = 2.0 * p.mu * np.asarray(energies)[:, None, None] * eye
    phi = ldmd.initialize_inelastic(W[0][None] - shift)
    psi = np.broadcast_to(eye, phi.shape).copy()
    for s, h in enumerate(hs):
        Ws, Wm, We = (W[2 * s + i][None] - shift for i in range(3))
        k1p, k1f = phi, Ws @ psi
        k2p, k2f = phi + h * k1f, Wm @ (psi + h * k1p)
        k3p, k3f = phi + h * k2f, Wm @ (psi + h * k2p)
        k4p, k4f = phi + 2 * h * k3f, We @ (psi + 2 * h * k3p)
        psi = psi + (h / 3) * (k1p + 2 * k2p + 2 * k3p + k4p)
        phi = phi + (h / 3) * (k1f + 2 * k2f + 2 * k3f + k4f)
        Q, _ = np.linalg.qr(np.concatenate([psi, phi], axis=1))
        psi, phi = Q[:, :N], Q[:, N:]
    return phi @ np.linalg.inv(psi)


def _smat(p, Y, energies):
    B, L = omatch.bmat_bf_to_sf(p.U, p.ladder, p.j_couple, p.Jtot)
    return omatch.get_smat(omatch.transform_bf_to_sf(Y, B), p.sectors[-1][1], energies, p.mu, p.E_int, L)


_SYSTEM = dict(jmax=4, Jtot=2, n_energy=3, boundaries=(3.0, 15.0), n_theta=12)  # N = 12, 50-500 cm^-1


@pytest.fixture(scope="module")
def rk4_limit():
    """Richardson limit of the RK4 S-matrices at half steps 0.005 and 0.0025 bohr."""
    energies = synthetic.atom_rotor(**_SYSTEM).energies
    S = []
    for h in (0.005, 0.0025):
        _, _, p = oracle_problem(synthetic.atom_rotor(**_SYSTEM, half_step=h))
        S.append(_smat(p, _rk4_log_derivative(p, energies), energies))
    return energies, [b + (b - a) / 15.0 for a, b in zip(*S)]


def _pipeline_errors(energies, limit, monkeypatch=None, variant=None):
    if variant is not None:
        monkeypatch.setattr(ldmd, "correction_matrices", _variant_corrections(variant))
    errs = []
    for h in (0.02, 0.01, 0.005):
        _, _, p = oracle_problem(synthetic.atom_rotor(**_SYSTEM, half_step=h))
        _, _, _, S = opipe.solve(p, energies)
        errs.append(max(np.abs(a - b).max() for a, b in zip(S, limit)))
    return errs


def test_pipeline_smatrix_converges_to_the_rk4_limit(rk4_limit):
    energies, limit = rk4_limit
    assert all(s.shape[0] > 1 for s in limit)
    for s in limit:
        assert omatch.unitarity_defect(s) < 1e-6
    errs = _pipeline_errors(energies, limit)
    assert np.all(_orders(errs) > 3.6), errs
    assert errs[-1] < 4e-5, errs


def test_planted_mistake_fails_the_pipeline_check(rk4_limit, monkeypatch):
    energies, limit = rk4_limit
    errs = _pipeline_errors(energies, limit, monkeypatch, "swap_start_end")
    assert not (np.all(_orders(errs) > 3.6) and errs[-1] < 4e-5), errs
