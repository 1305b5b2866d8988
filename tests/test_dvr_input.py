"""Monomer sine DVR / PODVR and the TOML front end (CPU).

Reference KATs re-expressed: tests/unit/test_dvr.py, test_podvr.py (shapes,
orthonormality, j = 0 states equal the vibrational ones, argument checks) and the
parsing half of test_driver.py / test_input_auto_detection.py.  The DVR is pinned
beyond them by the closed-form sine-DVR kinetic matrix, the particle in a box
(exact in this basis) and the harmonic oscillator.
"""

from pathlib import Path

import numpy as np
import pytest

import pyticc_amd as ticc
from pyticc_amd.basis.dvr import (
    build_RovibDVR,
    build_RovibPODVR,
    build_SineDVR,
    build_VibPODVR,
    phase_fix,
    sine_dvr_kinetic,
)
from pyticc_amd.input.driver import _resolve_scattering_type, build_system
from pyticc_amd.system import ScatteringType


def _closed_form_kinetic(a, b, n, mass):
    # Colbert & Miller, J. Chem. Phys. 96, 1982 (1992), sine basis on (a, b)
    m = n + 1
    i = np.arange(1, n + 1)[:, None]
    j = np.arange(1, n + 1)[None, :]
    with np.errstate(divide="ignore"):
        off = (-1.0) ** (i - j) * (1.0 / np.sin(np.pi * (i - j) / (2 * m)) ** 2 - 1.0 / np.sin(np.pi * (i + j) / (2 * m)) ** 2)
    T = np.where(i == j, (2.0 * m * m + 1.0) / 3.0 - 1.0 / np.sin(np.pi * i / m) ** 2, off)
    return np.pi**2 / (4.0 * mass * (b - a) ** 2) * T


@pytest.mark.parametrize("n", [2, 7, 24])
def test_kinetic_matrix_equals_the_closed_form(n):
    T = sine_dvr_kinetic(1.0, 4.0, n, 900.0)
    ref = _closed_form_kinetic(1.0, 4.0, n, 900.0)
    np.testing.assert_allclose(T, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_particle_in_a_box_is_exact():
    a, b, n, mass = 0.0, 2.5, 16, 3.0
    dvr = build_SineDVR(a, b, n, mass, lambda r: np.zeros_like(r))
    k = np.pi * np.arange(1, n + 1) / (b - a)
    np.testing.assert_allclose(dvr.eigen_val, k * k / (2 * mass), rtol=1e-12)


def test_harmonic_oscillator_levels():
    mass, kf = 900.0, 3.0
    dvr = build_SineDVR(0.0, 4.0, 100, mass, lambda r: 0.5 * kf * (r - 2.0) ** 2)
    w = np.sqrt(kf / mass)
    np.testing.assert_allclose(dvr.eigen_val[:5], w * (np.arange(5) + 0.5), rtol=1e-10)


def test_phase_convention():
    v = phase_fix(np.array([5e-5, -0.3, 0.2]))
    np.testing.assert_array_equal(v, [-5e-5, 0.3, -0.2])
    M = phase_fix(np.array([[-1.0, 2.0], [3.0, -4.0]]))
    np.testing.assert_array_equal(M, [[1.0, 2.0], [-3.0, -4.0]])
    with pytest.raises(ValueError, match="below the threshold"):
        phase_fix(np.array([1e-5, -1e-6]))


def test_primitive_and_contracted_rovib_bases():
    # reference tests/unit/test_dvr.py:9-41
    dvr = build_SineDVR(1.0, 3.0, 12, 500.0, lambda r: 0.02 * (r - 2.0) ** 2)
    assert isinstance(build_RovibDVR(dvr, vmax=1, jmax=1, mass=500.0), ticc.RovibBasis)
    assert isinstance(build_RovibPODVR(dvr, n_podvr=5, vmax=1, jmax=1, mass=500.0), ticc.RovibBasis)

    dvr = build_SineDVR(1.0, 4.0, 24, 900.0, lambda r: 0.03 * (r - 2.0) ** 2)
    rovib = build_RovibDVR(dvr, vmax=1, jmax=2, mass=900.0)
    assert rovib.grids is dvr.grids
    assert rovib.E_vj.shape == (2, 3) and rovib.WF_vj.shape == (24, 2, 3)
    np.testing.assert_allclose(rovib.E_vj[:, 0], dvr.eigen_val[:2], atol=1e-13)
    np.testing.assert_allclose(rovib.WF_vj[:, :, 0], dvr.eigen_vec[:, :2], atol=1e-12)
    for j in range(3):
        np.testing.assert_allclose(rovib.WF_vj[:, :, j].T @ rovib.WF_vj[:, :, j], np.eye(2), atol=1e-13)

    dvr0 = build_SineDVR(1.0, 3.0, 8, 500.0, lambda r: np.zeros_like(r))
    with pytest.raises(ValueError, match="non-negative"):
        build_RovibDVR(dvr0, vmax=-1, jmax=0, mass=500.0)
    with pytest.raises(ValueError, match="positive and finite"):
        build_RovibDVR(dvr0, vmax=0, jmax=0, mass=0.0)


def test_podvr():
    # reference tests/unit/test_podvr.py
    dvr = build_SineDVR(1.0, 3.0, 16, 1000.0, lambda x: 0.01 * (x - 2.0) ** 2)
    vib = build_VibPODVR(dvr, n_podvr=4, vmax=2)
    assert vib.grids.shape == (4,) and vib.energies.shape == (3,) and vib.wavefunctions.shape == (4, 3)
    np.testing.assert_allclose(vib.wavefunctions.T @ vib.wavefunctions, np.eye(3), atol=1e-13)
    # the contracted basis holds the exact lowest states: their energies are kept
    np.testing.assert_allclose(vib.energies, dvr.eigen_val[:3], rtol=1e-12)
    # a complete contraction reproduces the primitive rovibrational energies (same spectrum)
    full = build_RovibPODVR(dvr, n_podvr=16, vmax=3, jmax=2, mass=1000.0)
    prim = build_RovibDVR(dvr, vmax=3, jmax=2, mass=1000.0)
    np.testing.assert_allclose(full.E_vj, prim.E_vj, rtol=1e-11)
    # PODVR points are the eigenvalues of r in the contracted basis: inside (a, b), ascending
    assert np.all(np.diff(full.grids) > 0) and full.grids[0] > 1.0 and full.grids[-1] < 3.0


_AD_INPUT = """
type = "A+BC"
atom = "Ar"
diatom = ["H", "F"]
Jtot = 1
system_parity = -1
energies_cm = "energies.dat"

[approximation]
method = "cs"

[basis]
r = [1.5, 4.5]
n_dvr = 20
n_podvr = 1
vmax = 0
jmax = 1

[quadrature]
n_theta = 4

[channels]
vmin_Y = 0
exchange_parity_Y = 0
E_Y_cut_cm = 1000.0
K_cut = "none"

[potential_grid]
radial_boundaries = [3.0, 3.2]
radial_half_steps = [0.1]

[propagation]
mode = "inelastic"
"""

_DD_INPUT = """
type = "AB+CD"
diatom_X = ["H", "H"]
diatom_Y = ["H", "F"]
Jtot = 0
system_parity = 1
energies_cm = [100.0]

[approximation]
method = "nncc"
K_delta = 1

[basis_X]
r = [0.4, 3.5]
n_dvr = 20
n_podvr = 1
vmax = 0
jmax = 0

[basis_Y]
r = [0.7, 4.7]
n_dvr = 20
n_podvr = 1
vmax = 0
jmax = 0

[quadrature]
n_theta_X = 3
n_theta_Y = 3
n_phi = 4

[channels]
vmin_X = 0
vmin_Y = 0
exchange_parity_X = 1
exchange_parity_Y = 0
E_X_cut_cm = 1000.0
E_Y_cut_cm = 1000.0
K_cut = "none"

[potential_grid]
radial_boundaries = [3.0, 3.2]
radial_half_steps = [0.1]

[propagation]
mode = "inelastic"
"""


def _zero_pes(two=False):
    zero = lambda r: np.zeros_like(r)  # noqa: E731
    return ticc.PESWrapper(interaction=lambda RR, c: np.zeros(c.shape[1]), monomer_X=zero if two else None,
                           monomer_Y=zero)


def _parse(text):
    from pyticc_amd.input.driver import tomllib

    return tomllib.loads(text)


def test_parse_atom_diatom_input():
    # the parsing half of the reference's test_driver.py:9-60
    config = _parse(_AD_INPUT)
    assert _resolve_scattering_type(config) is ScatteringType.ATOM_DIATOM
    system, quad = build_system(config, ScatteringType.ATOM_DIATOM, _zero_pes())
    assert system.basis.n_channel == 3 and quad == {"n_theta": 4}
    assert system.approx is ticc.Approx.CS
    assert any(c.K == 1 for c in system.basis)
    assert len(ticc.build_cs_blocks(system.basis)) == 2


def test_parse_diatom_diatom_input():
    # the parsing half of the reference's test_driver.py:203-278
    config = _parse(_DD_INPUT)
    assert _resolve_scattering_type(config) is ScatteringType.DIATOM_DIATOM
    system, quad = build_system(config, ScatteringType.DIATOM_DIATOM, _zero_pes(two=True))
    assert system.basis.n_channel == 1 and system.approx is ticc.Approx.NNCC
    assert quad == {"n_theta_X": 3, "n_theta_Y": 3, "n_phi": 4}


def test_type_detection_and_refusals():
    assert _resolve_scattering_type({"atom": "Ar", "diatom": ["H", "F"]}) is ScatteringType.ATOM_DIATOM
    assert _resolve_scattering_type({"diatom_X": ["H", "H"], "diatom_Y": ["H", "F"]}) is ScatteringType.DIATOM_DIATOM
    with pytest.raises(ValueError, match="Cannot infer"):
        _resolve_scattering_type({"atom": "Ar"})
    with pytest.raises(ValueError, match="Unsupported TOML type"):
        _resolve_scattering_type({"type": "A+BCD"})
    for refused in ({"type": "A+BC_electric"}, {"type": "A+BC_diabatic"}, {"atom": "Ar", "electric": {}},
                    {"fine_structure": {"atom": {}, "diatom": {}}}):
        with pytest.raises(NotImplementedError):
            _resolve_scattering_type(refused)


def test_input_errors(tmp_path: Path):
    bad = _AD_INPUT.replace('K_cut = "none"', "K_cut = -2")
    with pytest.raises(ValueError, match="K_cut"):
        build_system(_parse(bad), ScatteringType.ATOM_DIATOM, _zero_pes())
    with pytest.raises(ValueError, match="Missing required input 'Jtot'"):
        build_system(_parse(_AD_INPUT.replace("Jtot = 1\n", "")), ScatteringType.ATOM_DIATOM, _zero_pes())
    with pytest.raises(ValueError, match="monomer_Y"):
        build_system(_parse(_AD_INPUT), ScatteringType.ATOM_DIATOM, ticc.PESWrapper(interaction=lambda R, c: 0))
    f = tmp_path / "input.toml"
    f.write_text(_AD_INPUT + '\n[pes]\npath = "."\nsources = ["interaction-PES.f"]\n')
    np.savetxt(tmp_path / "energies.dat", [100.0])
    with pytest.raises(NotImplementedError, match="library"):
        ticc.run(f)
    f.write_text("type = [")
    with pytest.raises(ValueError, match="Failed to read"):
        ticc.run(f)
