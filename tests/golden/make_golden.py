"""Generate the committed golden fixtures (run on the CPU; minutes for the N=800 case).

    python tests/golden/make_golden.py [bessel] [cg] [atom_rotor_N66] [rotor_rotor_N800]
    python tests/golden/make_golden.py --parts DIR [--only 4,15]   (the N = 800 energies split
                                                                    across processes, cached in DIR)

Fixtures (NumPy .npz, no pickled objects):
  bessel_mpmath.npz   Riccati-Bessel j, n, j', n' (j = sqrt(pi x/2) J_{l+1/2}(x)) and the
                      K log-derivative K'_nu/K_nu at integer and non-integer orders,
                      evaluated with mpmath at 40 significant digits -- an independent
                      high-precision check of the scipy-based oracle and of the device code.
  cg_sympy.npz        Clebsch-Gordan coefficients from sympy.physics.wigner (exact
                      rationals under a square root, rounded once to FP64).
  smat_atom_rotor_N66.npz    oracle S-matrices of the atom + rigid-rotor system (N = 66, J = 10,
                      p = +1, 1350 sectors 3-30 bohr) at 5 of its 20 energies.
  smat_rotor_rotor_N800.npz    oracle S-matrices of the two-rotor system (N = 800, J = 4, p = -1,
                      2000 sectors 4-40 bohr, 15 x 15 x 12 grid) at 6 of its 64 energies:
                      every element of S at energies 4, 15 and 31 (17, 202 and 473 open
                      channels; the upper triangle, S being symmetric), the diagonal and 64
                      seeded random columns at energies 47 and 63 (729 and 800 open), none
                      open at energy 0; per energy the oracle's unitarity defect and the FP64
                      noise floor -- the larger max |dS| between the solve-based oracle and
                      the same oracle with its sector inverses formed another way: explicitly
                      (LAPACK getri, oracle.ldmd.sector_step_inverse) or by blocked
                      Gauss-Jordan with partial pivoting (oracle.ldmd.sector_step_gj), the
                      device's algorithm class, or that together with W contracted in
                      another summation order (oracle.contraction.contract_split), as the
                      device changes both; all are equally valid FP64 orderings; and
                      the matching step's own floor, the distance between the last
                      variant's Y matched with solves and with explicit inverses
                      (oracle.matching.get_smat(inverse=True)).

Every fixture is written by the oracle (tests/../oracle) or by mpmath/sympy; nothing
here reads the reference sources.
"""

from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

CONFIG1_ENERGIES = (0, 4, 9, 14, 19)
CONFIG2_ENERGIES = (0, 4, 15, 31, 47, 63)
CONFIG2_FULL_MAX_OPEN = 480  # energies with at most this many open channels store all of S
CONFIG2_SAMPLE_COLUMNS = 64  # else: the diagonal and this many seeded random columns


def make_bessel() -> None:
    import mpmath as mp

    mp.mp.dps = 40
    ell = [0.0, 0.0, 1.0, 2.3, 3.0, 4.2, 10.0, 17.0, 20.3, 1.3, 0.7, 12.5, 30.0, 5.0, 0.5, 24.0]
    x = [1.7, 0.05, 3.0, 4.2, 516.0, 9.0, 1.5, 60.0, 300.0, 2500.0, 1.99, 35.0, 45.0, 0.3, 1000.0, 31.0]
    rows = []
    for l, xv in zip(ell, x):
        nu = mp.mpf(l) + mp.mpf(1) / 2
        X = mp.mpf(xv)
        s = mp.sqrt(mp.pi * X / 2)
        J, Y = mp.besselj(nu, X), mp.bessely(nu, X)
        Jp, Yp = mp.diff(lambda t: mp.besselj(nu, t), X), mp.diff(lambda t: mp.bessely(nu, t), X)
        rows.append([float(s * J), float(s * Y), float(s * (Jp + J / (2 * X))), float(s * (Yp + Y / (2 * X)))])
    nu_k = [0.5, 1.7, 1.5, 3.5, 10.5, 48.5, 0.5, 20.5, 2.7, 7.5]
    x_k = [0.3, 3.1, 1.999, 2.0, 7.0, 2500.0, 100.0, 0.05, 476.0, 40.0]
    klogd = []
    for nu, xv in zip(nu_k, x_k):
        f = lambda t: mp.besselk(nu, t)  # noqa: E731
        klogd.append(float(mp.diff(f, xv) / f(xv)))
    np.savez(HERE / "bessel_mpmath.npz", ell=np.array(ell), x=np.array(x), riccati=np.array(rows),
             nu=np.array(nu_k), xk=np.array(x_k), k_logd=np.array(klogd))


def make_cg() -> None:
    from sympy.physics.wigner import clebsch_gordan

    rng = np.random.default_rng(11)
    cases = []
    while len(cases) < 200:
        j1, j2 = (int(v) for v in rng.integers(0, 9, 2))
        J = int(rng.integers(abs(j1 - j2), j1 + j2 + 1))
        m1 = int(rng.integers(-j1, j1 + 1))
        m2 = int(rng.integers(-j2, j2 + 1))
        if abs(m1 + m2) > J:
            continue
        val = float(clebsch_gordan(j1, j2, J, m1, m2, m1 + m2).evalf(30))
        cases.append((j1, m1, j2, m2, J, val))
    arr = np.array(cases)
    np.savez(HERE / "cg_sympy.npz", idx=arr[:, :5].astype(np.int64), value=arr[:, 5])


def _oracle_smat(workload, energy_index):
    from oracle import matching as omatch
    from oracle import pipeline as opipe
    from tests.helpers import oracle_problem

    _, _, prob = oracle_problem(workload)
    E = workload.energies[list(energy_index)]
    _, _, _, S = opipe.solve(prob, E)
    return E, S, [omatch.unitarity_defect(s) if s.size else 0.0 for s in S]


def make_atom_rotor_N66() -> None:
    from pyticc_amd import synthetic

    w = synthetic.atom_rotor()
    E, S, defect = _oracle_smat(w, CONFIG1_ENERGIES)
    out = {"energy_index": np.array(CONFIG1_ENERGIES), "energies": E, "defect": np.array(defect)}
    for i, s in enumerate(S):
        out[f"S{i}"] = s
    np.savez_compressed(HERE / "smat_atom_rotor_N66.npz", **out)


def _rotor_rotor_N800_part(w, prob, k: int, parts_dir):
    """S of both FP64 orderings at energy index k (cached in parts_dir when given)."""
    from oracle import ldmd
    from oracle import pipeline as opipe

    path = None if parts_dir is None else Path(parts_dir) / f"n800_energy{k}.npz"
    if path is not None and path.exists():
        d = np.load(path)
        return d["S"], d["S_inv"]
    E = w.energies[[k]]
    Y_solve, Y_inv = opipe.propagate(prob, E, steps=(ldmd.sector_step, ldmd.sector_step_inverse))
    S = opipe.match(prob, Y_solve, E)[2][0]
    S_inv = opipe.match(prob, Y_inv, E)[2][0]
    if path is not None:
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, S=S, S_inv=S_inv)
    return S, S_inv


def _rotor_rotor_N800_gj_part(w, prob, k: int, parts_dir):
    """S of the blocked Gauss-Jordan ordering at energy index k (cached in parts_dir)."""
    from oracle import ldmd
    from oracle import pipeline as opipe

    path = None if parts_dir is None else Path(parts_dir) / f"n800_energy{k}_gj.npz"
    if path is not None and path.exists():
        return np.load(path)["S_gj"]
    E = w.energies[[k]]
    Y_gj = opipe.propagate(prob, E, steps=(ldmd.sector_step_gj,))
    S_gj = opipe.match(prob, Y_gj, E)[2][0]
    if path is not None:
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, S_gj=S_gj)
    return S_gj


def _rotor_rotor_N800_alt_part(w, prob, k: int, parts_dir):
    """S at energy index k with both of the solve's FP64 orderings changed at once:
    W contracted with each angular sum in two halves (oracle.contraction.contract_split)
    and the sector inverses by blocked Gauss-Jordan (cached in parts_dir)."""
    from oracle import contraction as ocontract
    from oracle import ldmd
    from oracle import pipeline as opipe

    path = None if parts_dir is None else Path(parts_dir) / f"n800_energy{k}_alt.npz"
    if path is not None and path.exists():
        d = np.load(path)
        return d["S_alt"], d["S_alt_match_inv"]
    E = w.energies[[k]]
    Y = opipe.propagate(prob, E, steps=(ldmd.sector_step_gj,), contract=ocontract.contract_split)
    S_alt = opipe.match(prob, Y, E)[2][0]
    # the same Y matched with explicit inverses: the matching step's own floor
    S_alt_match_inv = opipe.match(prob, Y, E, inverse=True)[2][0]
    if path is not None:
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, S_alt=S_alt, S_alt_match_inv=S_alt_match_inv)
    return S_alt, S_alt_match_inv


def make_rotor_rotor_N800(parts_dir=None, only=None) -> None:
    """The N = 800 fixture, one energy at a time; with parts_dir each energy's S
    matrices are cached there (so separate processes can split the energies, and
    an interrupted run resumes), and with `only` just those energy indices are
    computed and nothing is assembled."""
    from oracle import matching as omatch
    from pyticc_amd import synthetic
    from tests.helpers import oracle_problem

    w = synthetic.rotor_rotor()
    _, _, prob = oracle_problem(w)
    todo = CONFIG2_ENERGIES if only is None else tuple(only)
    parts = {k: _rotor_rotor_N800_part(w, prob, k, parts_dir) for k in todo}
    gj = {k: _rotor_rotor_N800_gj_part(w, prob, k, parts_dir) for k in todo}
    alt = {k: _rotor_rotor_N800_alt_part(w, prob, k, parts_dir) for k in todo}
    if only is not None:
        return
    S = [parts[k][0] for k in CONFIG2_ENERGIES]
    S_inv = [parts[k][1] for k in CONFIG2_ENERGIES]
    S_gj = [gj[k] for k in CONFIG2_ENERGIES]
    dist = lambda a, b: float(np.max(np.abs(a - b))) if a.size else 0.0  # noqa: E731
    floor_inv = np.array([dist(a, b) for a, b in zip(S, S_inv)])
    floor_gj = np.array([dist(a, b) for a, b in zip(S, S_gj)])
    floor_alt = np.array([dist(a, alt[k][0]) for a, k in zip(S, CONFIG2_ENERGIES)])
    floor_match = np.array([dist(alt[k][0], alt[k][1]) for k in CONFIG2_ENERGIES])
    rng = np.random.default_rng(800)
    out = {"energy_index": np.array(CONFIG2_ENERGIES), "energies": w.energies[list(CONFIG2_ENERGIES)],
           "defect": np.array([omatch.unitarity_defect(s) if s.size else 0.0 for s in S]),
           "floor_inverse": floor_inv, "floor_gj": floor_gj, "floor_alt": floor_alt,
           "floor": np.maximum(np.maximum(floor_inv, floor_gj), floor_alt), "floor_match": floor_match,
           "n_open": np.array([s.shape[0] for s in S])}
    for i, s in enumerate(S):
        n = s.shape[0]
        if n <= CONFIG2_FULL_MAX_OPEN:
            iu = np.triu_indices(n)
            out[f"S{i}_upper"] = s[iu]
        else:
            cols = np.sort(rng.choice(n, CONFIG2_SAMPLE_COLUMNS, replace=False))
            out[f"S{i}_colidx"] = cols
            out[f"S{i}_cols"] = s[:, cols]
            out[f"S{i}_diag"] = np.diag(s).copy()
    np.savez_compressed(HERE / "smat_rotor_rotor_N800.npz", **out)

if __name__ == "__main__":
    args = sys.argv[1:]
    parts_dir = None
    if "--parts" in args:  # --parts DIR [--only k1,k2]: cache (and split) the N = 800 energies
        i = args.index("--parts")
        parts_dir = args[i + 1]
        del args[i : i + 2]
    only = None
    if "--only" in args:
        i = args.index("--only")
        only = [int(k) for k in args[i + 1].split(",")]
        del args[i : i + 2]
    if parts_dir is not None or only is not None:
        t = time.perf_counter()
        make_rotor_rotor_N800(parts_dir, only)
        print(f"rotor_rotor_N800 {only or 'all'}: {time.perf_counter() - t:.1f} s", flush=True)
        sys.exit(0)
    which = args or ["bessel", "cg", "atom_rotor_N66", "rotor_rotor_N800"]
    for name in which:
        t = time.perf_counter()
        {"bessel": make_bessel, "cg": make_cg, "atom_rotor_N66": make_atom_rotor_N66, "rotor_rotor_N800": make_rotor_rotor_N800}[name]()
        print(f"{name}: {time.perf_counter() - t:.1f} s", flush=True)
