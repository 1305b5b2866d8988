"""HIP path against the committed golden fixtures: mpmath Bessel tables and oracle
S-matrices of the atom + rigid-rotor system (N = 66, all 1350 sectors) and of the
two-rotor benchmark system at its full size (N = 800, 2000 sectors).

Tolerances at full size come from FP64 noise floors recorded with the fixture: the
largest distance between the oracle and its other equally valid orderings after
2000 sectors at N = 800 (sector inverses formed explicitly, by blocked
Gauss-Jordan, and that together with a different contraction summation order) for
the contraction and sweep, and the distance between two orderings of the matching
on the same Y for the matching.  The device's S is split accordingly: its own Y
matched by the oracle is held to the sweep's floor, and its own matching of that Y
to the matching's floor (each: the 1e-10 bar or twice the floor, whichever is
larger).  The reduced systems use 1e-10."""

from pathlib import Path

import numpy as np
import pytest

import pyticc_amd as ticc
from oracle import matching as omatch
from pyticc_amd import synthetic
from pyticc_amd.match.bessel import k_logd_batch, riccati_bessel_batch
from tests.helpers import hamiltonian_problem, n800_match_tolerance, n800_tolerance, rotor_rotor_N800_reference

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
TOL_S = 1.0e-10


def test_device_bessel_matches_mpmath(gpu_device):
    d = np.load(GOLDEN / "bessel_mpmath.npz")
    got = riccati_bessel_batch(d["ell"], d["x"], gpu_device.index)
    for g, ref in zip(got, d["riccati"]):
        env = max(np.hypot(ref[0], ref[1]), 1e-300)
        envp = max(np.hypot(ref[2], ref[3]), 1e-300)
        np.testing.assert_allclose(g[:2], ref[:2], rtol=0, atol=2e-13 * env)
        np.testing.assert_allclose(g[2:], ref[2:], rtol=0, atol=2e-13 * envp)
    np.testing.assert_allclose(k_logd_batch(d["nu"], d["xk"], gpu_device.index), d["k_logd"], rtol=1e-13)


def test_solve_atom_rotor_N66_matches_fixture(gpu_device):
    d = np.load(GOLDEN / "smat_atom_rotor_N66.npz")
    w = synthetic.atom_rotor()
    grid = ticc.prepare_potential(w.system, w.boundaries, w.half_steps, **w.quadrature)
    res = ticc.solve(w.system, w.energies, grid, ticc.Propagation(device="gpu"))
    worst = 0.0
    for i, e in enumerate(d["energy_index"]):
        ref = d[f"S{i}"]
        assert res.Smat[e].shape == ref.shape
        worst = max(worst, float(np.max(np.abs(res.Smat[e] - ref))))
    assert worst <= TOL_S, worst
    assert max(omatch.unitarity_defect(s) for s in res.Smat if s.size) < 1e-10
    # the S-matrix text table carries the same elements, S[out, in], incoming outermost
    e = int(d["energy_index"][0])
    lines = ticc.report.smatrix(res, energy_indices=e).splitlines()[2:]
    vals = np.array([complex(float(x.split()[-2]), float(x.split()[-1])) for x in lines])
    np.testing.assert_allclose(vals, res.Smat[e].T.ravel(), rtol=1e-15, atol=1e-16)


def test_solve_rotor_rotor_N800_full_size_matches_fixture(gpu_device):
    from oracle import pipeline as opipe
    from pyticc_amd.scattering.hamiltonian import build_hamiltonian

    path = GOLDEN / "smat_rotor_rotor_N800.npz"
    d = np.load(path)
    w = synthetic.rotor_rotor()
    idx = d["energy_index"]
    E = w.energies[idx]
    grid = ticc.prepare_potential(w.system, w.boundaries, w.half_steps, **w.quadrature)
    print("N = 800 full-size solve on the device", flush=True)
    res = ticc.solve(w.system, E, grid, ticc.Propagation(device="gpu", memory_mb=4096.0))
    # the device's propagated Y matched by the oracle: separates the sweep's
    # distance from the fixture from the matching's
    prob = hamiltonian_problem(build_hamiltonian(w.system, grid), grid)
    S_om = opipe.match(prob, np.asarray(res.Y_propagated), E)[2]
    report = []
    for i in range(idx.size):
        S = res.Smat[i]
        assert S.shape[0] == int(d["n_open"][i])
        if S.size == 0:  # no open channel at this energy
            continue
        r, c, ref = rotor_rotor_N800_reference(d, i)
        sweep = float(np.max(np.abs(S_om[i][r, c] - ref)))
        match = float(np.max(np.abs(S - S_om[i])))
        total = float(np.max(np.abs(S[r, c] - ref)))
        report.append((int(idx[i]), S.shape[0], total, sweep, n800_tolerance(d, i), match, n800_match_tolerance(d, i)))
        # the device S is as unitary as the oracle's on the same inputs
        assert omatch.unitarity_defect(S) <= 10.0 * float(d["defect"][i]) + 1e-12
    print("N = 800 two-rotor full size (energy, n_open, max|dS| total, sweep part, its tol, matching part, its tol):",
          report)
    for e, n, total, sweep, tol_s, match, tol_m in report:
        assert sweep <= tol_s, ("sweep", e, n, sweep, tol_s)
        assert match <= tol_m, ("matching", e, n, match, tol_m)
