"""Test oracle (CPU): an independent restatement of the setup arrays the LDMD oracle
consumes -- the channel list, the body-fixed centrifugal matrix U, the weighted
angular rows B_K and the PES on the propagation grid -- for atom + rigid rotor and
rigid rotor + rigid rotor systems.  Only tests/ may import it.

Restates (from the published definitions, in this module's own numerics):
  * channels: ``build_ChannelBasis`` (src/pyticc/basis/channel.py:491-576): every
    (state_X, state_Y, j12 in |jX - jY|..jX + jY, K in 0..min(j12, J, K_cut)); a K = 0
    channel only when parity (-1)^(jX + jY + j12 + J) = +1 (channel.py:428-437); a
    stable sort by the internal energy E_X + E_Y; rigid-rotor levels E = B j (j + 1)
    taken in ascending j;
  * U: ``get_Umat_BF`` (src/pyticc/matrix/centrifugal.py:27-85): J(J+1) + j12(j12+1)
    - 2K^2 on the diagonal and -c_K lambda+(J, K) lambda+(j12, K) between K and K + 1
    of one (state_X, state_Y, j12) ladder, c_0 = sqrt 2, c_K = 1 otherwise,
    lambda+(j, K) = sqrt(j(j+1) - K(K+1));
  * B_K rows: ``prepare`` of src/pyticc/matrix/interaction/atom_diatom.py:351-386
    (sqrt(w) N_jK(cos theta)) and diatom_diatom.py:16-137 (sum over omega_X of
    CG(jX omega_X, jY K-omega_X | j12 K) sqrt(w_X) N_{jX omega_X} sqrt(w_Y) N_{jY K-omega_X}
    sqrt(w_phi) (cos, sin)(omega_X phi), normalisation 1/pi), with N_jK(x) =
    sqrt(2 pi) Y_jK(theta, 0) from scipy's spherical harmonics (Condon-Shortley
    phase), Clebsch-Gordan coefficients from sympy, and Gauss-Legendre rules from
    numpy.polynomial.legendre.leggauss;
  * the PES grid: ``prepare_potential`` (src/pyticc/scattering/potential.py:182-222)
    with ``get_Vgrid_atom_diatom`` / ``get_Vgrid_diatom_diatom`` (src/pyticc/pes/adiabatic.py:110-143, 224-256): the PES at every
    ordered radial point on the meshgrid(ij) of (r, theta) or (r_X, r_Y, theta_X,
    theta_Y, phi), theta = arccos of the Gauss-Legendre nodes, phi on [0, pi].

Pinned by tests/test_oracle_setup.py: the channel counts the reference's tests
state (N = 66 / 55 for the atom + rotor system, 800 with K blocks 80/205/192/173/150
for the two-rotor system), U of a two-channel ladder by hand, the rows' quadrature
orthonormality, and the 3j/9j matrix elements of tests/test_angular_closed_forms.py;
then the product's arrays are compared with these.
"""

from __future__ import annotations

from functools import lru_cache

import numpy as np
from scipy.special import sph_harm_y

from oracle import grid as ogrid


def gl_rule(lower: float, upper: float, n: int, symmetry: bool = False):
    """Gauss-Legendre nodes and weights on [lower, upper]; symmetry keeps the first
    half of a 2n-point rule with doubled weights."""
    m = 2 * n if symmetry else n
    x, w = np.polynomial.legendre.leggauss(m)
    c, h = 0.5 * (upper + lower), 0.5 * (upper - lower)
    if symmetry:
        return c + h * x[:n], 2.0 * h * w[:n]
    return c + h * x, h * w


def rigid_levels(B: float, jmax: int):
    """[(j, E_j)] of a rigid rotor, E_j = B j (j + 1), ascending j."""
    return [(j, B * j * (j + 1)) for j in range(jmax + 1)]


def channel_list(levels_X, levels_Y, Jtot: int, parity: int, K_cut: int | None = None):
    """[(jX, jY, j12, K, E_int)] in the basis order: enumeration (X level, Y level,
    j12, K), K = 0 parity rule, stable sort by E_int."""
    out = []
    for jX, EX in levels_X:
        for jY, EY in levels_Y:
            for j12 in range(abs(jX - jY), jX + jY + 1):
                kmax = min(j12, Jtot) if K_cut is None else min(j12, Jtot, K_cut)
                for K in range(kmax + 1):
                    if K == 0 and parity * (-1) ** (jX + jY + j12 + Jtot) != 1:
                        continue
                    out.append((jX, jY, j12, K, float(EX + EY)))
    return sorted(out, key=lambda c: c[4])  # Python's sort is stable


def lambda_plus(j: int, K: int) -> float:
    return float(np.sqrt(j * (j + 1) - K * (K + 1)))


def umat_bf(chans, Jtot: int) -> np.ndarray:
    """Dimensionless body-fixed centrifugal matrix of a channel list."""
    n = len(chans)
    U = np.zeros((n, n))
    where = {}
    for a, (jX, jY, j12, K, _) in enumerate(chans):
        U[a, a] = Jtot * (Jtot + 1) + j12 * (j12 + 1) - 2 * K * K
        where[(jX, jY, j12, K)] = a
    for (jX, jY, j12, K), a in where.items():
        b = where.get((jX, jY, j12, K + 1))
        if b is None:
            continue
        c = -(np.sqrt(2.0) if K == 0 else 1.0) * lambda_plus(Jtot, K) * lambda_plus(j12, K)
        U[a, b] = U[b, a] = c
    return U


def norm_yjk(j: int, K: int, x) -> np.ndarray:
    """sqrt(2 pi) Y_jK(arccos x, 0): the normalised associated Legendre function
    (int_{-1}^{1} N_jK^2 dx = 1), Condon-Shortley phase."""
    return np.sqrt(2.0 * np.pi) * sph_harm_y(j, K, np.arccos(np.asarray(x, dtype=np.float64)), 0.0).real


@lru_cache(maxsize=None)
def _cg(j1: int, m1: int, j2: int, m2: int, J: int) -> float:
    from sympy.physics.wigner import clebsch_gordan

    return float(clebsch_gordan(j1, j2, J, m1, m2, m1 + m2))


def rows_atom_diatom(chans, ct, wt):
    """{K: (channel indices, B_real rows (n_K, n_theta))} of an atom + rigid rotor basis
    (the rotor is the Y monomer; its j is chans' jY)."""
    out = {}
    sw = np.sqrt(wt)
    for K in sorted({c[3] for c in chans}):
        idx = [a for a, c in enumerate(chans) if c[3] == K]
        out[K] = (np.array(idx), np.stack([sw * norm_yjk(chans[a][1], K, ct) for a in idx]))
    return out


def rows_diatom_diatom(chans, ctX, wX, ctY, wY, phi, wphi, phi_of_Y: bool = False):
    """{K: (channel indices, B_real, B_imag)} of a rigid rotor + rigid rotor basis,
    rows over the (theta_X, theta_Y, phi) grid in C order.  phi_of_Y plants a
    mistake (the torsion phase taken from omega_Y instead of omega_X) for the
    tests that show the comparison catches it."""
    out = {}
    sX, sY, sP = np.sqrt(wX), np.sqrt(wY), np.sqrt(wphi)
    for K in sorted({c[3] for c in chans}):
        idx = [a for a, c in enumerate(chans) if c[3] == K]
        re = np.zeros((len(idx), ctX.size * ctY.size * phi.size))
        im = np.zeros_like(re)
        for r, a in enumerate(idx):
            jX, jY, j12 = chans[a][:3]
            R = np.zeros((ctX.size, ctY.size, phi.size))
            I = np.zeros_like(R)
            for wx in range(-jX, jX + 1):
                wy = K - wx
                if abs(wy) > jY:
                    continue
                amp = _cg(jX, wx, jY, wy, j12) * np.multiply.outer(sX * norm_yjk(jX, wx, ctX), sY * norm_yjk(jY, wy, ctY))
                w = wy if phi_of_Y else wx
                R += amp[:, :, None] * (sP * np.cos(w * phi))[None, None, :]
                I += amp[:, :, None] * (sP * np.sin(w * phi))[None, None, :]
            re[r] = R.reshape(-1)
            im[r] = I.reshape(-1)
        out[K] = (np.array(idx), re, im)
    return out


def radial_points(boundaries, half_steps) -> np.ndarray:
    return ogrid.radial_points(ogrid.build_radial_sectors(boundaries, half_steps))


def pes_grid_atom_diatom(pes_many, R, r, ct) -> np.ndarray:
    """V (n_R, n_r * n_theta) on the meshgrid(ij) of (r, theta = arccos ct)."""
    g = np.meshgrid(np.asarray(r, dtype=np.float64), np.arccos(ct), indexing="ij")
    coords = np.stack([x.reshape(-1) for x in g])
    return np.asarray(pes_many(np.asarray(R, dtype=np.float64), coords), dtype=np.float64)


def pes_grid_diatom_diatom(pes_many, R, rX, rY, ctX, ctY, phi) -> np.ndarray:
    """V (n_R, n_grid) on the meshgrid(ij) of (r_X, r_Y, theta_X, theta_Y, phi)."""
    g = np.meshgrid(np.asarray(rX, dtype=np.float64), np.asarray(rY, dtype=np.float64), np.arccos(ctX),
                    np.arccos(ctY), np.asarray(phi, dtype=np.float64), indexing="ij")
    coords = np.stack([x.reshape(-1) for x in g])
    return np.asarray(pes_many(np.asarray(R, dtype=np.float64), coords), dtype=np.float64)
