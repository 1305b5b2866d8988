"""The weighted basis rows B_K (vmat.prepare of the reference,
src/pyticc/matrix/interaction/atom_diatom.py:351-386 and diatom_diatom.py:16-137)
pinned by angular-momentum closed forms, independently of the product's own
associated Legendre functions and Clebsch-Gordan coefficients.

A PES that is one angular term is contracted through the product's rows (with
the oracle's contraction, common.py:49-62) and compared with the matrix element
given by Wigner 3j / 9j symbols (sympy) of the standard body-fixed channel
functions |(j_X j_Y) j12 K> = sum CG(j_X w, j_Y K-w | j12 K) Y_{j_X w} Y_{j_Y K-w}:

  A + BC, V = P_lam(cos theta):
    <j K|P_lam|j' K> = (-1)^K sqrt((2j+1)(2j'+1)) (j lam j'; 0 0 0) (j lam j'; -K 0 K);

  AB + CD, V = [Y^{l1}(theta_X, phi) x Y^{l2}(theta_Y, 0)]^{lam}_0 (scipy's spherical
  harmonics, Condon-Shortley phase; real and even in phi when l1 + l2 + lam is even):
    <(jX jY) j12 K|V|(jX' jY') j12' K> = (-1)^{j12-K} (j12 lam j12'; -K 0 K)
        sqrt((2 j12+1)(2 j12'+1)(2 lam+1)) {jX jX' l1; jY jY' l2; j12 j12' lam}
        <jX||Y^{l1}||jX'> <jY||Y^{l2}||jY'>,
    <j||Y^l||j'> = (-1)^j sqrt((2j+1)(2l+1)(2j'+1) / (4 pi)) (j l j'; 0 0 0)
  (Edmonds 5.4.1, 5.4.5, 7.1.5).

A slip that keeps the rows orthonormal -- a phi phase taken from the wrong rotor, a
K < 0 sign, a swapped CG order -- passes the constant-V identity test but not these.
"""

from functools import lru_cache

import numpy as np
import pytest
from scipy.special import eval_legendre, sph_harm_y
from sympy.physics.wigner import clebsch_gordan as sym_cg
from sympy.physics.wigner import wigner_3j, wigner_9j

from oracle import contraction as ocon
from pyticc_amd import AtomSpec, build_ScattSystem, prepare_potential, rigid_rotor
from pyticc_amd.constants import AMU2AU, CM2AU
from pyticc_amd.pes import PESWrapper
from pyticc_amd.scattering.hamiltonian import build_hamiltonian


@lru_cache(maxsize=None)
def _w3j(a, b, c, d, e, f):
    return float(wigner_3j(a, b, c, d, e, f))


@lru_cache(maxsize=None)
def _w9j(*args):
    return float(wigner_9j(*args, prec=40))


def _contracted(system, quad, pes):
    system = build_ScattSystem(*system[0], potential=PESWrapper(interaction=lambda R, c: pes(np.array([R]), c)[0], interaction_many=pes), **system[1])
    grid = prepare_potential(system, (5.0, 5.1), (0.05,), **quad)
    ham = build_hamiltonian(system, grid)
    vb = ham.vbasis
    blocks = [(np.asarray(vb.channel_indices[K]), vb.B_real[K], None if vb.B_imag is None else vb.B_imag[K])
              for K in sorted(vb.channel_indices)]
    V = grid.values.reshape(grid.radial_points.size, -1)[:1]
    return ham.basis, ocon.contract(blocks, ham.basis.n_channel, vb.normalization, V)[0]


@pytest.mark.parametrize("lam", [0, 1, 2, 3, 5])
@pytest.mark.parametrize("parity", [1, -1])
def test_atom_diatom_rows_give_gaunt_matrix_elements(lam, parity):
    def pes(R, coords):
        return np.broadcast_to(eval_legendre(lam, np.cos(coords[-1])), (np.size(R), coords.shape[1])).copy()

    basis, V = _contracted(((AtomSpec(), rigid_rotor(10.0 * CM2AU, 6)),
                            dict(Jtot=4, system_parity=parity, reduced_mass=10.0 * AMU2AU)), {"n_theta": 24}, pes)
    ref = np.zeros_like(V)
    for a, ca in enumerate(basis):
        for b, cb in enumerate(basis):
            if ca.K != cb.K:
                continue
            j, jp, K = ca.mis_X.j + ca.mis_Y.j, cb.mis_X.j + cb.mis_Y.j, ca.K  # the atom has j = 0
            ref[a, b] = (-1) ** K * np.sqrt((2 * j + 1) * (2 * jp + 1)) * _w3j(j, lam, jp, 0, 0, 0) * _w3j(j, lam, jp, -K, 0, K)
    assert basis.n_channel > 10
    np.testing.assert_allclose(V, ref, rtol=0, atol=1e-13)


def _reduced_Y(j, l, jp):
    return (-1) ** j * np.sqrt((2 * j + 1) * (2 * l + 1) * (2 * jp + 1) / (4 * np.pi)) * _w3j(j, l, jp, 0, 0, 0)


def _bispherical(l1, l2, lam):
    """[Y^{l1}(theta_X, phi) x Y^{l2}(theta_Y, 0)]^{lam}_0 on the PES coordinate rows."""
    def pes(R, coords):
        thX, thY, phi = coords[2], coords[3], coords[4]
        v = np.zeros(coords.shape[1], dtype=np.complex128)
        for m in range(-min(l1, l2), min(l1, l2) + 1):
            cg = float(sym_cg(l1, l2, lam, m, -m, 0))
            if cg:
                v += cg * sph_harm_y(l1, m, thX, phi) * sph_harm_y(l2, -m, thY, 0.0 * phi)
        assert np.max(np.abs(v.imag)) < 1e-14
        return np.broadcast_to(v.real, (np.size(R), v.size)).copy()
    return pes


@pytest.mark.parametrize("terms", [(1, 1, 0), (1, 1, 2), (2, 2, 2), (2, 1, 3), (2, 0, 2), (0, 2, 2), (2, 2, 4)])
@pytest.mark.parametrize("parity", [1, -1])
def test_diatom_diatom_rows_give_bispherical_matrix_elements(terms, parity):
    l1, l2, lam = terms
    basis, V = _contracted(((rigid_rotor(10.0 * CM2AU, 2), rigid_rotor(7.3 * CM2AU, 3)),
                            dict(Jtot=3, system_parity=parity, reduced_mass=10.0 * AMU2AU)),
                           {"n_theta_X": 10, "n_theta_Y": 10, "n_phi": 24}, _bispherical(l1, l2, lam))
    ref = np.zeros_like(V)
    for a, ca in enumerate(basis):
        for b, cb in enumerate(basis):
            if ca.K != cb.K:
                continue
            jX, jY, j12, K = ca.mis_X.j, ca.mis_Y.j, ca.j_couple, ca.K
            jXp, jYp, j12p = cb.mis_X.j, cb.mis_Y.j, cb.j_couple
            nine = _w9j(jX, jXp, l1, jY, jYp, l2, j12, j12p, lam)
            ref[a, b] = ((-1) ** (j12 - K) * _w3j(j12, lam, j12p, -K, 0, K)
                         * np.sqrt((2 * j12 + 1) * (2 * j12p + 1) * (2 * lam + 1)) * nine
                         * _reduced_Y(jX, l1, jXp) * _reduced_Y(jY, l2, jYp))
    assert basis.n_channel > 20 and np.max(np.abs(ref)) > 0.05
    np.testing.assert_allclose(V, ref, rtol=0, atol=1e-13)


# -- the oracle's own rows (oracle/setup.py) against the same closed forms ----------------

from oracle import setup as osetup  # noqa: E402


def _oracle_contracted_dd(parity, pes):
    ch = osetup.channel_list(osetup.rigid_levels(10.0 * CM2AU, 2), osetup.rigid_levels(7.3 * CM2AU, 3), 3, parity)
    ctX, wX = osetup.gl_rule(-1.0, 1.0, 10)
    ctY, wY = osetup.gl_rule(-1.0, 1.0, 10)
    phi, wphi = osetup.gl_rule(0.0, np.pi, 24)
    rows = osetup.rows_diatom_diatom(ch, ctX, wX, ctY, wY, phi, wphi)
    g = np.meshgrid([1.0], [1.0], np.arccos(ctX), np.arccos(ctY), phi, indexing="ij")
    V = pes(np.array([5.0]), np.stack([x.reshape(-1) for x in g]))
    blocks = [(idx, re, im) for _, (idx, re, im) in sorted(rows.items())]
    return ch, ocon.contract(blocks, len(ch), 1.0 / np.pi, V)[0]


@pytest.mark.parametrize("terms", [(1, 1, 2), (2, 1, 3), (2, 2, 4)])
@pytest.mark.parametrize("parity", [1, -1])
def test_oracle_diatom_diatom_rows_give_bispherical_matrix_elements(terms, parity):
    l1, l2, lam = terms
    ch, V = _oracle_contracted_dd(parity, _bispherical(l1, l2, lam))
    ref = np.zeros_like(V)
    for a, (jX, jY, j12, K, _) in enumerate(ch):
        for b, (jXp, jYp, j12p, Kp, _) in enumerate(ch):
            if K != Kp:
                continue
            ref[a, b] = ((-1) ** (j12 - K) * _w3j(j12, lam, j12p, -K, 0, K)
                         * np.sqrt((2 * j12 + 1) * (2 * j12p + 1) * (2 * lam + 1))
                         * _w9j(jX, jXp, l1, jY, jYp, l2, j12, j12p, lam)
                         * _reduced_Y(jX, l1, jXp) * _reduced_Y(jY, l2, jYp))
    assert np.max(np.abs(ref)) > 0.05
    np.testing.assert_allclose(V, ref, rtol=0, atol=1e-13)


@pytest.mark.parametrize("lam", [1, 2, 5])
def test_oracle_atom_diatom_rows_give_gaunt_matrix_elements(lam):
    ch = osetup.channel_list([(0, 0.0)], osetup.rigid_levels(10.0 * CM2AU, 6), 4, 1)
    ct, wt = osetup.gl_rule(-1.0, 1.0, 24)
    rows = osetup.rows_atom_diatom(ch, ct, wt)
    V = eval_legendre(lam, ct)[None, :]
    V = ocon.contract([(idx, re, None) for _, (idx, re) in sorted(rows.items())], len(ch), 1.0, V)[0]
    ref = np.zeros_like(V)
    for a, (_, j, _, K, _) in enumerate(ch):
        for b, (_, jp, _, Kp, _) in enumerate(ch):
            if K == Kp:
                ref[a, b] = (-1) ** K * np.sqrt((2 * j + 1) * (2 * jp + 1)) * _w3j(j, lam, jp, 0, 0, 0) * _w3j(j, lam, jp, -K, 0, K)
    np.testing.assert_allclose(V, ref, rtol=0, atol=1e-13)
