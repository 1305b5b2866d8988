"""Oracle and host setup against the committed golden fixtures (CPU).

The fixtures come from tests/golden/make_golden.py: mpmath (40 digits) and sympy
tables, and oracle S-matrices of the atom + rigid-rotor system (N = 66).
"""

from pathlib import Path

import numpy as np

from oracle import bessel as obessel
from oracle import matching as omatch
from oracle import pipeline as opipe
from pyticc_amd import synthetic
from pyticc_amd.basis.angle import clebsch_gordan
from tests.helpers import oracle_problem

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_oracle_riccati_bessel_matches_mpmath():
    d = np.load(GOLDEN / "bessel_mpmath.npz")
    for l, x, ref in zip(d["ell"], d["x"], d["riccati"]):
        got = np.array(obessel.riccati_bessel_jy(float(l), float(x)))
        env = max(np.hypot(ref[0], ref[1]), 1e-300)
        envp = max(np.hypot(ref[2], ref[3]), 1e-300)
        np.testing.assert_allclose(got[:2], ref[:2], rtol=0, atol=2e-13 * env)
        np.testing.assert_allclose(got[2:], ref[2:], rtol=0, atol=2e-13 * envp)


def test_oracle_k_log_derivative_matches_mpmath():
    d = np.load(GOLDEN / "bessel_mpmath.npz")
    got = [obessel.modified_bessel_K_logD(float(n), float(x)) for n, x in zip(d["nu"], d["xk"])]
    np.testing.assert_allclose(got, d["k_logd"], rtol=1e-13)


def test_clebsch_gordan_matches_sympy():
    d = np.load(GOLDEN / "cg_sympy.npz")
    got = np.array([clebsch_gordan(int(j1), int(m1), int(j2), int(m2), int(J)) for j1, m1, j2, m2, J in d["idx"]])
    np.testing.assert_allclose(got, d["value"], rtol=0, atol=1e-14)


def test_oracle_reproduces_atom_rotor_N66_fixture():
    d = np.load(GOLDEN / "smat_atom_rotor_N66.npz")
    w = synthetic.atom_rotor()
    _, _, prob = oracle_problem(w)
    idx = d["energy_index"]
    np.testing.assert_array_equal(w.energies[idx], d["energies"])
    _, _, _, S = opipe.solve(prob, w.energies[idx])
    for i, s in enumerate(S):
        ref = d[f"S{i}"]
        assert s.shape == ref.shape
        np.testing.assert_allclose(s, ref, rtol=0, atol=1e-12)
        assert omatch.unitarity_defect(s) < 1e-12
