"""Partial cross sections of the product (pyticc_amd.crosssection) against the
oracle's textbook restatement and the closed forms that pin it: the single-channel
phase-shift result, detailed balance, and the level sums of |delta - S|^2."""

import numpy as np
import pytest

import pyticc_amd as ticc
from oracle import matching as om
from pyticc_amd import synthetic
from pyticc_amd.crosssection import degeneracy, level_of
from pyticc_amd.result import ScatteringResult


def _random_symmetric_unitary(rng, n):
    K = rng.standard_normal((n, n))
    K = K + K.T
    I = np.eye(n)
    return np.linalg.solve(I + 1j * K, I - 1j * K)


def _result_with_random_S(seed=5):
    w = synthetic.atom_rotor(jmax=6, Jtot=4, n_energy=3)
    basis = w.system.basis
    rng = np.random.default_rng(seed)
    S = tuple(_random_symmetric_unitary(rng, int(np.sum(basis.E_int < E))) for E in w.energies)
    mu = w.system.reduced_mass
    res = ScatteringResult(basis=basis, Etot=w.energies, Y_propagated=None, asymptotic_transform=None, L=None, Smat=S,
                           reduced_mass=mu)
    return res, basis, mu


def test_cross_sections_match_oracle_restatement():
    res, basis, mu = _result_with_random_S()
    got = ticc.partial_cross_sections(res)
    assert len(got) == len(res.Smat)
    for cs, E, S in zip(got, res.Etot, res.Smat):
        open_idx = np.flatnonzero(basis.E_int < E)
        labels = [level_of(basis[i]) for i in open_idx]
        key = {lv: n for n, lv in enumerate(cs.levels)}
        lab = np.array([key[lv] for lv in labels])
        k = np.sqrt(2 * mu * (E - basis.E_int[open_idx]))
        g = {key[lv]: degeneracy(lv) for lv in cs.levels}
        levs, sig = om.partial_cross_sections(S, lab, k, g, basis.Jtot)
        assert list(levs) == list(range(len(cs.levels)))
        np.testing.assert_allclose(cs.sigma, sig, rtol=1e-13, atol=0)
        assert np.all(np.diff(cs.E_level) >= 0)


def test_cross_sections_detailed_balance_and_level_sums():
    res, basis, mu = _result_with_random_S(seed=8)
    for cs, E, S in zip(ticc.partial_cross_sections(res), res.Etot, res.Smat):
        k2 = 2 * mu * (E - cs.E_level)
        g = np.array([degeneracy(lv) for lv in cs.levels], dtype=float)
        w = (k2 * g)[:, None] * cs.sigma
        np.testing.assert_allclose(w, w.T, rtol=1e-12, atol=1e-12 * np.abs(w).max())
        # sum over all (a, b) of k_a^2 g_a sigma(a -> b) / (pi (2J+1)) = sum |I - S|^2
        total = float(np.sum(np.abs(np.eye(S.shape[0]) - S) ** 2))
        np.testing.assert_allclose(w.sum() / (np.pi * (2 * basis.Jtot + 1)), total, rtol=1e-12)


def test_cross_section_single_channel_phase_shift():
    w = synthetic.atom_rotor(jmax=0, Jtot=0, n_energy=1)
    basis = w.system.basis
    assert basis.n_channel == 1
    delta = 0.37
    res = ScatteringResult(basis=basis, Etot=w.energies, Y_propagated=None, asymptotic_transform=None, L=None,
                           Smat=(np.array([[np.exp(2j * delta)]]),), reduced_mass=w.system.reduced_mass)
    (cs,) = res.partial_cross_sections()
    k2 = 2 * w.system.reduced_mass * (w.energies[0] - basis.E_int[0])
    np.testing.assert_allclose(cs.sigma[0, 0], 4 * np.pi / k2 * np.sin(delta) ** 2, rtol=1e-14)


def test_cross_sections_validate_inputs():
    res, _, _ = _result_with_random_S()
    with pytest.raises(ValueError):
        ticc.partial_cross_sections(res, reduced_mass=-1.0)
    bad = ScatteringResult(basis=res.basis, Etot=res.Etot, Y_propagated=None, asymptotic_transform=None, L=None,
                           Smat=(np.eye(2),) * len(res.Etot), reduced_mass=res.reduced_mass)
    with pytest.raises(ValueError):
        ticc.partial_cross_sections(bad)
