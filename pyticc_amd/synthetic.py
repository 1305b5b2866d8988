"""Seeded synthetic exact-CC workloads (rigid rotors, Tang-Toennies PES).

atom_rotor:   A + rigid rotor, V(R, theta) = sum_{l<=6} C_l(R) P_l(cos theta),
              C_l(R) = A_l exp(-b_l R) - f6(b_l R) D_l / R^6  (Tang-Toennies damping),
              A_l in [5, 50] Eh, b_l in [1.5, 2.5] / bohr, D_l in [1, 20] Eh bohr^6.
rotor_rotor:  two rigid rotors, V = sum_{l1,l2<=6, m<=min(l1,l2)} C_{l1 l2 m}(R)
              Pbar_{l1}^m(cos theta_X) Pbar_{l2}^m(cos theta_Y) cos(m phi), same radial form.
The default sizes are the benchmark workloads: N = 66 (j<=10, J=10, p=+1) and
N = 800 (j_X<=4, j_Y<=9, J=4, p=-1); the rotor-rotor energy grid 10.87-1000.87 cm^-1
stays >= 0.44 cm^-1 away from every channel threshold.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.special import eval_legendre, gammaln, lpmv

from pyticc_amd.basis.monomer import AtomSpec, rigid_rotor
from pyticc_amd.constants import AMU2AU, CM2AU
from pyticc_amd.pes import PESWrapper
from pyticc_amd.system import Approx, build_ScattSystem


def tang_toennies_f6(x: np.ndarray) -> np.ndarray:
    """f6(x) = 1 - exp(-x) sum_{k<=6} x^k / k!."""
    s = np.zeros_like(x)
    term = np.ones_like(x)
    for k in range(7):
        if k:
            term = term * x / k
        s = s + term
    return 1.0 - np.exp(-x) * s


def _radial(R: np.ndarray, A: np.ndarray, b: np.ndarray, D: np.ndarray) -> np.ndarray:
    """(n_R, n_term) Tang-Toennies radial coefficients."""
    R = np.asarray(R, dtype=np.float64)[:, None]
    x = b[None, :] * R
    return A[None, :] * np.exp(-x) - tang_toennies_f6(x) * D[None, :] / R**6


@dataclass(frozen=True)
class AtomRotorPES:
    A: np.ndarray
    b: np.ndarray
    D: np.ndarray

    @classmethod
    def seeded(cls, seed: int, lmax: int = 6) -> "AtomRotorPES":
        rng = np.random.default_rng(seed)
        n = lmax + 1
        return cls(rng.uniform(5.0, 50.0, n), rng.uniform(1.5, 2.5, n), rng.uniform(1.0, 20.0, n))

    def interaction_many(self, R, coords):
        theta = coords[1]
        ang = np.stack([eval_legendre(l, np.cos(theta)) for l in range(self.A.size)])  # (n_l, n_grid)
        return _radial(np.atleast_1d(R), self.A, self.b, self.D) @ ang

    def interaction(self, R, coords):
        return self.interaction_many(np.array([R]), coords)[0]

    def wrapper(self) -> PESWrapper:
        return PESWrapper(interaction=self.interaction, interaction_many=self.interaction_many)


def _pbar(l: int, m: int, x: np.ndarray) -> np.ndarray:
    f = np.exp(0.5 * (np.log((2.0 * l + 1.0) / 2.0) + gammaln(l - m + 1.0) - gammaln(l + m + 1.0)))
    return f * lpmv(m, l, x)


@dataclass(frozen=True)
class RotorRotorPES:
    terms: tuple  # (l1, l2, m)
    A: np.ndarray
    b: np.ndarray
    D: np.ndarray

    @classmethod
    def seeded(cls, seed: int, lmax: int = 6) -> "RotorRotorPES":
        terms = tuple((l1, l2, m) for l1 in range(lmax + 1) for l2 in range(lmax + 1) for m in range(min(l1, l2) + 1))
        rng = np.random.default_rng(seed)
        n = len(terms)
        # isotropic term strongest; anisotropic amplitudes shrink with l1 + l2
        scale = np.asarray([1.0 / (1.0 + l1 + l2) for l1, l2, _ in terms])
        return cls(terms, rng.uniform(5.0, 50.0, n) * scale, rng.uniform(1.5, 2.5, n), rng.uniform(1.0, 20.0, n) * scale)

    def angular(self, coords) -> np.ndarray:
        ctX, ctY, phi = np.cos(coords[2]), np.cos(coords[3]), coords[4]
        return np.stack([_pbar(l1, m, ctX) * _pbar(l2, m, ctY) * np.cos(m * phi) for l1, l2, m in self.terms])

    def interaction_many(self, R, coords):
        return _radial(np.atleast_1d(R), self.A, self.b, self.D) @ self.angular(coords)

    def interaction(self, R, coords):
        return self.interaction_many(np.array([R]), coords)[0]

    def wrapper(self) -> PESWrapper:
        return PESWrapper(interaction=self.interaction, interaction_many=self.interaction_many)


@dataclass(frozen=True)
class Workload:
    """A prepared system plus its radial grid, quadrature options and energies (atomic units)."""

    name: str
    system: object
    boundaries: tuple
    half_steps: tuple
    quadrature: dict
    energies: np.ndarray


def atom_rotor(seed: int = 1, *, jmax: int = 10, Jtot: int = 10, parity: int = 1, n_energy: int = 20,
               boundaries=(3.0, 30.0), half_step: float = 0.01, n_theta: int = 20, B_rot_cm: float = 10.0,
               mu_amu: float = 10.0, e_range_cm=(50.0, 500.0), approx: Approx = Approx.EXACT, K_delta: int = 1) -> Workload:
    """Atom + rigid rotor workload (default: N = 66, 20 energies, 1350 sectors)."""
    system = build_ScattSystem(AtomSpec(), rigid_rotor(B_rot_cm * CM2AU, jmax), Jtot=Jtot, system_parity=parity,
                               approx=approx, K_delta=K_delta, potential=AtomRotorPES.seeded(seed).wrapper(),
                               reduced_mass=mu_amu * AMU2AU)
    E = np.linspace(*e_range_cm, n_energy) * CM2AU
    return Workload("atom_rotor", system, tuple(boundaries), (half_step,), {"n_theta": n_theta}, E)


def rotor_rotor(seed: int = 2, *, jmax_X: int = 4, jmax_Y: int = 9, Jtot: int = 4, parity: int = -1,
                n_energy: int = 64, boundaries=(4.0, 40.0), half_step: float = 0.009, n_theta_X: int = 15,
                n_theta_Y: int = 15, n_phi: int = 12, B_X_cm: float = 10.0, B_Y_cm: float = 7.3,
                mu_amu: float = 10.0, e_range_cm=(10.87, 1000.87), approx: Approx = Approx.EXACT,
                K_delta: int = 1) -> Workload:
    """Two rigid rotors (default: N = 800, 64 energies, 2000 sectors, 15 x 15 x 12 angular grid)."""
    system = build_ScattSystem(rigid_rotor(B_X_cm * CM2AU, jmax_X), rigid_rotor(B_Y_cm * CM2AU, jmax_Y), Jtot=Jtot,
                               system_parity=parity, approx=approx, K_delta=K_delta,
                               potential=RotorRotorPES.seeded(seed).wrapper(), reduced_mass=mu_amu * AMU2AU)
    E = np.linspace(*e_range_cm, n_energy) * CM2AU
    quad = {"n_theta_X": n_theta_X, "n_theta_Y": n_theta_Y, "n_phi": n_phi}
    return Workload("rotor_rotor", system, tuple(boundaries), (half_step,), quad, E)


@dataclass(frozen=True)
class BlockSpec:
    """One block of the batching workload: the n_channel lowest channels of the
    two-rotor (Jtot, parity) basis."""

    Jtot: int
    parity: int
    n_channel: int


def rotor_rotor_block_specs(seed: int = 3, *, n_block: int = 256, n_range=(40, 1200), Jmax: int = 8,
                            **rotor_kw) -> tuple:
    """Seeded blocks of uneven size: N log-uniform in n_range; each is the N lowest
    channels (by threshold energy) of a (J, parity) basis of the two-rotor system
    drawn uniformly among those with at least N channels (J <= Jmax)."""
    rng = np.random.default_rng(seed)
    lo, hi = n_range
    sizes = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), n_block))).astype(int)
    counts = {(J, p): rotor_rotor(Jtot=J, parity=p, n_energy=1, **rotor_kw).system.basis.n_channel
              for J in range(Jmax + 1) for p in (1, -1)}
    specs = []
    for n in sizes:
        fits = sorted(k for k, c in counts.items() if c >= n)
        if not fits:
            raise ValueError(f"no (J, parity) basis with J <= {Jmax} holds {n} channels")
        J, p = fits[int(rng.integers(len(fits)))]
        specs.append(BlockSpec(J, p, int(n)))
    return tuple(specs)


def rotor_rotor_blocks(specs, *, n_energy: int = 32, **rotor_kw) -> dict:
    """The two-rotor workload of every (J, parity) the block specs use (same PES,
    radial grid, quadrature and energies), keyed by (J, parity)."""
    keys = sorted({(s.Jtot, s.parity) for s in specs})
    return {k: rotor_rotor(Jtot=k[0], parity=k[1], n_energy=n_energy, **rotor_kw) for k in keys}
