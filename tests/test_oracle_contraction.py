"""Pins of the oracle contraction and W assembly (CPU): constant V -> V I
(reference tests/unit/test_interaction.py:116-151), block order, radial batch,
W signs (test_radial.py:6-24)."""

import numpy as np
from scipy.special import roots_legendre

from oracle import contraction as oc
from pyticc_amd import AtomSpec, build_ScattSystem, rigid_rotor
from pyticc_amd.basis.angle import gauss_legendre_dvr
from pyticc_amd.matrix.interaction import prepare_atom_diatom, prepare_diatom_diatom


def _blocks(vb):
    return [(np.asarray(vb.channel_indices[K]), vb.B_real[K], None if vb.B_imag is None else vb.B_imag[K]) for K in sorted(vb.channel_indices)]


def test_constant_potential_identity_atom_diatom():
    sysm = build_ScattSystem(AtomSpec(), rigid_rotor(1e-4, 3), Jtot=2, system_parity=1)
    ct, wt = roots_legendre(6)
    vb = prepare_atom_diatom(sysm.basis, sysm.monomer_Y.rovib, ct, wt)
    V = oc.contract(_blocks(vb), sysm.n_channel, vb.normalization, np.full((1, 6), 2.5))
    np.testing.assert_allclose(V[0], 2.5 * np.eye(sysm.n_channel), atol=1e-13)


def test_constant_potential_identity_diatom_diatom():
    sysm = build_ScattSystem(rigid_rotor(1e-4, 2), rigid_rotor(1e-4, 2), Jtot=1, system_parity=-1)
    ctX, wX = gauss_legendre_dvr(-1, 1, 5)
    ctY, wY = gauss_legendre_dvr(-1, 1, 5)
    phi, wp = gauss_legendre_dvr(0, np.pi, 16)
    vb = prepare_diatom_diatom(sysm.basis, sysm.monomer_X.rovib, sysm.monomer_Y.rovib, ctX, wX, ctY, wY, phi, wp)
    V = oc.contract(_blocks(vb), sysm.n_channel, vb.normalization, np.full((1, 5 * 5 * 16), 1.75))
    np.testing.assert_allclose(V[0], 1.75 * np.eye(sysm.n_channel), atol=1e-12)


def test_cross_K_elements_are_zero_and_radial_batch():
    sysm = build_ScattSystem(AtomSpec(), rigid_rotor(1e-4, 3), Jtot=3, system_parity=1)
    ct, wt = roots_legendre(7)
    vb = prepare_atom_diatom(sysm.basis, sysm.monomer_Y.rovib, ct, wt)
    rng = np.random.default_rng(1)
    V = rng.standard_normal((2, 7))
    out = oc.contract(_blocks(vb), sysm.n_channel, 1.0, V)
    K = np.asarray([c.K for c in sysm.basis])
    assert np.all(out[:, K[:, None] != K[None, :]] == 0.0)
    np.testing.assert_allclose(out[1], oc.contract(_blocks(vb), sysm.n_channel, 1.0, V[1:])[0])
    np.testing.assert_allclose(out, np.swapaxes(out, 1, 2))


def test_w_signs():
    U = np.array([[2.0, 0.5], [0.5, 6.0]])
    V = np.array([[0.1, 0.02], [0.02, -0.3]])
    W = oc.w_full(2.0, 0.7, 1.5, np.array([0.0, 0.2]), U, V)
    expect = U / 4.0 + 3.0 * V + np.diag(3.0 * (np.array([0.0, 0.2]) - 0.7))
    np.testing.assert_allclose(W, expect)


def test_split_summation_order_agrees():
    """contract_split (each angular sum in two halves) is the same arithmetic in
    another order: equal to rounding on random data, and bit for bit on small
    integers, where every partial sum is exact."""
    rng = np.random.default_rng(4)
    idx = np.arange(6)
    Br, Bi = rng.standard_normal((6, 41)), rng.standard_normal((6, 41))
    V = rng.standard_normal((3, 41))
    a = oc.contract([(idx, Br, Bi)], 6, 0.7, V)
    b = oc.contract_split([(idx, Br, Bi)], 6, 0.7, V)
    np.testing.assert_allclose(b, a, rtol=0, atol=1e-13 * np.abs(a).max())
    Bn, Vn = rng.integers(-3, 4, (6, 41)).astype(float), rng.integers(-5, 6, (2, 41)).astype(float)
    np.testing.assert_array_equal(oc.contract_split([(idx, Bn, None)], 6, 1.0, Vn),
                                  oc.contract([(idx, Bn, None)], 6, 1.0, Vn))
