"""The oracle's own statement of the setup (oracle/setup.py) against the product's:
channel lists, the centrifugal matrix U, the weighted angular rows B_K, the
quadrature and the PES grid of the benchmark systems (the N = 66 atom + rotor and
the N = 800 two-rotor system) and smaller variants.  The oracle side is pinned
first by what the reference's tests and the mathematics fix: the channel counts
(N = 66 / 55, 800 with K blocks 80/205/192/173/150), a Coriolis ladder by hand,
the rows' quadrature orthonormality (constant V -> V I), and a planted mistake
that the row comparison must catch.  With these equal, the GPU-vs-oracle S
comparisons, whose oracle problem takes the product's setup arrays
(tests/helpers.py), no longer rest on that setup unchecked."""

import numpy as np
import pytest

from oracle import setup as osetup
from pyticc_amd import prepare_potential, synthetic
from pyticc_amd.constants import CM2AU
from pyticc_amd.matrix.centrifugal import get_Umat_BF
from pyticc_amd.scattering.hamiltonian import build_hamiltonian


def _product_channels(basis):
    return [(c.mis_X.j, c.mis_Y.j, c.j_couple, c.K, c.E_int) for c in basis]


def _atom_rotor_oracle(jmax=10, Jtot=10, parity=1, B_cm=10.0):
    return osetup.channel_list([(0, 0.0)], osetup.rigid_levels(B_cm * CM2AU, jmax), Jtot, parity)


def _rotor_rotor_oracle(jmax_X=4, jmax_Y=9, Jtot=4, parity=-1, B_X=10.0, B_Y=7.3):
    return osetup.channel_list(osetup.rigid_levels(B_X * CM2AU, jmax_X), osetup.rigid_levels(B_Y * CM2AU, jmax_Y),
                               Jtot, parity)


def test_oracle_channel_counts_are_the_reference_counts():
    assert len(_atom_rotor_oracle()) == 66
    assert len(_atom_rotor_oracle(parity=-1)) == 55
    ch = _rotor_rotor_oracle()
    assert len(ch) == 800
    assert [sum(1 for c in ch if c[3] == K) for K in range(5)] == [80, 205, 192, 173, 150]
    E = [c[4] for c in ch]
    assert E == sorted(E)


def test_oracle_umat_of_a_coriolis_ladder():
    # J = 1, one ladder j12 = 1 with K = 0, 1 (the K = 0 member needs (-1)^(jX+jY+j12+J) = parity)
    ch = [(0, 1, 1, 0, 0.0), (0, 1, 1, 1, 0.0)]
    U = osetup.umat_bf(ch, 1)
    np.testing.assert_allclose(U, [[4.0, -2.0 * np.sqrt(2.0)], [-2.0 * np.sqrt(2.0), 2.0]], rtol=0, atol=1e-15)
    # different ladders (other j12) do not couple
    U2 = osetup.umat_bf([(1, 1, 1, 1, 0.0), (1, 1, 2, 2, 0.0)], 3)
    assert U2[0, 1] == 0.0


@pytest.mark.parametrize("case", ["atom_rotor_N66", "atom_rotor_N55", "rotor_rotor_small", "rotor_rotor_N800"])
def test_product_channels_and_umat_equal_the_oracle(case):
    if case.startswith("atom"):
        parity = 1 if case.endswith("66") else -1
        w = synthetic.atom_rotor(parity=parity, n_energy=1)
        ch = _atom_rotor_oracle(parity=parity)
    elif case == "rotor_rotor_small":
        w = synthetic.rotor_rotor(jmax_X=2, jmax_Y=3, Jtot=3, parity=1, n_energy=1)
        ch = _rotor_rotor_oracle(jmax_X=2, jmax_Y=3, Jtot=3, parity=1)
    else:
        w = synthetic.rotor_rotor(n_energy=1)
        ch = _rotor_rotor_oracle()
    basis = w.system.basis
    assert _product_channels(basis) == ch  # order, quantum numbers and E_int exactly
    np.testing.assert_allclose(get_Umat_BF(basis), osetup.umat_bf(ch, basis.Jtot), rtol=0, atol=1e-13)


def test_oracle_rows_are_orthonormal_under_the_quadrature():
    ctX, wX = osetup.gl_rule(-1.0, 1.0, 6)
    ctY, wY = osetup.gl_rule(-1.0, 1.0, 7)
    phi, wphi = osetup.gl_rule(0.0, np.pi, 24)  # enough nodes for the trigonometric products
    ch = osetup.channel_list(osetup.rigid_levels(1.0, 2), osetup.rigid_levels(1.3, 2), 3, -1)
    for K, (idx, re, im) in osetup.rows_diatom_diatom(ch, ctX, wX, ctY, wY, phi, wphi).items():
        np.testing.assert_allclose((re @ re.T + im @ im.T) / np.pi, np.eye(idx.size), atol=1e-13)
    ct, wt = osetup.gl_rule(-1.0, 1.0, 12)
    ch = osetup.channel_list([(0, 0.0)], osetup.rigid_levels(1.0, 6), 4, 1)
    for K, (idx, re) in osetup.rows_atom_diatom(ch, ct, wt).items():
        np.testing.assert_allclose(re @ re.T, np.eye(idx.size), atol=1e-13)


def _grid_and_rows(w):
    grid = prepare_potential(w.system, w.boundaries, w.half_steps, **w.quadrature)
    return grid, build_hamiltonian(w.system, grid).vbasis


@pytest.mark.parametrize("case", ["atom_rotor_N66", "rotor_rotor_small", "rotor_rotor_N800"])
def test_product_rows_quadrature_and_pes_grid_equal_the_oracle(case):
    if case == "atom_rotor_N66":
        w = synthetic.atom_rotor(n_energy=1)
        ch = _atom_rotor_oracle()
    elif case == "rotor_rotor_small":
        w = synthetic.rotor_rotor(jmax_X=2, jmax_Y=3, Jtot=3, parity=1, n_energy=1, boundaries=(4.0, 5.0))
        ch = _rotor_rotor_oracle(jmax_X=2, jmax_Y=3, Jtot=3, parity=1)
    else:
        w = synthetic.rotor_rotor(n_energy=1)
        ch = _rotor_rotor_oracle()
    grid, vb = _grid_and_rows(w)
    R = osetup.radial_points(w.boundaries, w.half_steps)
    np.testing.assert_array_equal(grid.radial_points, R)
    pes = w.system.potential
    if case.startswith("atom"):
        ct, wt = osetup.gl_rule(-1.0, 1.0, w.quadrature["n_theta"])
        np.testing.assert_allclose(grid.coordinate("cos_theta"), ct, rtol=0, atol=2e-15)
        np.testing.assert_allclose(grid.weight("theta"), wt, rtol=0, atol=2e-15)
        rows = {K: (idx, re, None) for K, (idx, re) in osetup.rows_atom_diatom(ch, ct, wt).items()}
        V = osetup.pes_grid_atom_diatom(pes.interaction_many, R, w.system.monomer_Y.rovib.grids, ct)
        norm = 1.0
    else:
        q = w.quadrature
        ctX, wX = osetup.gl_rule(-1.0, 1.0, q["n_theta_X"])
        ctY, wY = osetup.gl_rule(-1.0, 1.0, q["n_theta_Y"])
        phi, wphi = osetup.gl_rule(0.0, np.pi, q["n_phi"])
        for name, x in (("cos_theta_X", ctX), ("cos_theta_Y", ctY), ("phi", phi)):
            np.testing.assert_allclose(grid.coordinate(name), x, rtol=0, atol=4e-15)
        rows = osetup.rows_diatom_diatom(ch, ctX, wX, ctY, wY, phi, wphi)
        V = osetup.pes_grid_diatom_diatom(pes.interaction_many, R, w.system.monomer_X.rovib.grids,
                                          w.system.monomer_Y.rovib.grids, ctX, ctY, phi)
        norm = 1.0 / np.pi
    assert vb.normalization == pytest.approx(norm, rel=1e-15)
    assert sorted(vb.channel_indices) == sorted(rows)
    for K, (idx, re, im) in rows.items():
        np.testing.assert_array_equal(np.asarray(vb.channel_indices[K]), idx)
        scale = max(np.abs(re).max(), 0.0 if im is None else np.abs(im).max())  # one part can vanish
        np.testing.assert_allclose(vb.B_real[K], re, rtol=0, atol=1e-13 * scale)
        if im is not None:
            np.testing.assert_allclose(vb.B_imag[K], im, rtol=0, atol=1e-13 * scale)
    # the quadrature nodes agree to a few ulps, so V does to a few ulps of its scale
    np.testing.assert_allclose(grid.values.reshape(R.size, -1), V, rtol=1e-12, atol=1e-14 * np.abs(V).max())


def test_row_comparison_catches_a_planted_phase_mistake():
    """The torsion phase taken from omega_Y instead of omega_X keeps the rows
    orthonormal, so the constant-V identity cannot see it; the row comparison does."""
    w = synthetic.rotor_rotor(jmax_X=2, jmax_Y=3, Jtot=3, parity=1, n_energy=1, boundaries=(4.0, 5.0))
    ch = _rotor_rotor_oracle(jmax_X=2, jmax_Y=3, Jtot=3, parity=1)
    _, vb = _grid_and_rows(w)
    q = w.quadrature
    ctX, wX = osetup.gl_rule(-1.0, 1.0, q["n_theta_X"])
    ctY, wY = osetup.gl_rule(-1.0, 1.0, q["n_theta_Y"])
    phi, wphi = osetup.gl_rule(0.0, np.pi, q["n_phi"])
    bad = osetup.rows_diatom_diatom(ch, ctX, wX, ctY, wY, phi, wphi, phi_of_Y=True)
    worst = max(float(np.abs(vb.B_imag[K] - im).max()) for K, (_, _, im) in bad.items())
    assert worst > 1e-2
    for K, (idx, re, im) in bad.items():
        np.testing.assert_allclose((re @ re.T + im @ im.T) / np.pi, np.eye(idx.size), atol=1e-12)
