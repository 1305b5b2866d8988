"""The TOML front end end to end on the device: the reference's driver KATs
(tests/unit/test_driver.py:9-60, 203-278) and a vibrating-diatom A+BC input whose
compiled PES is named in the file, against the oracle solve of the same system."""

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import pyticc_amd as ticc
from oracle import pipeline
from pyticc_amd.input.driver import _resolve_scattering_type, build_system, tomllib
from pyticc_amd.scattering.hamiltonian import build_hamiltonian
from tests.helpers import hamiltonian_problem
from tests.test_dvr_input import _AD_INPUT, _DD_INPUT, _zero_pes
from tests.test_native_pes import C_SOURCE, HEADER

pytestmark = pytest.mark.gpu


def test_run_reads_compact_atom_diatom_input(gpu_device, tmp_path: Path):
    np.savetxt(tmp_path / "energies.dat", [100.0, 200.0])
    f = tmp_path / "input.toml"
    f.write_text(_AD_INPUT)
    result = ticc.run(f, pes=_zero_pes())
    assert result.basis.n_channel == 3
    assert isinstance(result, ticc.CoupledStatesResult)
    assert result.approx is ticc.Approx.CS
    assert any(c.K == 1 for c in result.basis)
    np.testing.assert_allclose(result.Etot * ticc.AU2CM, [100.0, 200.0])
    assert len(result.blocks) == 2
    assert result.timing is not None and result.timing.wall_seconds >= 0.0 and result.timing.cpu_seconds >= 0.0
    for block in result.blocks:
        for S in block.Smat_asymptotic:
            np.testing.assert_allclose(S.conj().T @ S, np.eye(S.shape[0]), atol=1e-12)


def test_run_reads_diatom_diatom_input(gpu_device, tmp_path: Path):
    f = tmp_path / "input.toml"
    f.write_text(_DD_INPUT)
    result = ticc.run(f, pes=_zero_pes(two=True))
    assert result.basis.n_channel == 1
    assert isinstance(result, ticc.CoupledStatesResult)
    assert result.approx is ticc.Approx.NNCC
    assert result.Etot.shape == (1,)
    assert len(result.blocks) == 1


_VIB_INPUT = """
atom = "Ar"
diatom = ["H", "F"]
Jtot = 2
system_parity = 1
energies_cm = [150.0, 420.0]

[pes]
library = "libpes.so"
monomer_Y = "pes_monomer"

[basis]
r = [0.9, 2.3]
n_dvr = 30
n_podvr = 4
vmax = 1
jmax = 5

[quadrature]
n_theta = 10

[channels]
E_Y_cut_cm = 5000.0
K_cut = "none"

[potential_grid]
radial_boundaries = [3.0, 7.0]
radial_half_steps = [0.04]

[propagation]
mode = "inelastic"
device = "gpu"
"""


def test_run_with_compiled_pes_matches_oracle(gpu_device, tmp_path: Path):
    """A vibrating diatom (sine DVR + PODVR, v <= 1) and a gcc-built PES named in
    [pes]: S from run() against the oracle solve of the same system and grid."""
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    (tmp_path / "pes.c").write_text(C_SOURCE)
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC", f"-I{HEADER.parent}", str(tmp_path / "pes.c"), "-o",
                    str(tmp_path / "libpes.so"), "-lm"], check=True)
    f = tmp_path / "input.toml"
    f.write_text(_VIB_INPUT)
    result = ticc.run(f)
    assert isinstance(result, ticc.ScatteringResult)
    assert {c.mis_Y.v for c in result.basis} == {0, 1}

    config = tomllib.loads(_VIB_INPUT)
    pes = ticc.load_native_pes(tmp_path / "libpes.so", monomer_Y="pes_monomer")
    system, quad = build_system(config, _resolve_scattering_type(config), pes)
    assert system.basis.channels == result.basis.channels
    grid = ticc.prepare_potential(system, (3.0, 7.0), (0.04,), **quad)
    prob = hamiltonian_problem(build_hamiltonian(system, grid), grid)
    _, _, _, S_ref = pipeline.solve(prob, result.Etot)
    for S, R in zip(result.Smat, S_ref):
        assert S.shape == R.shape and S.shape[0] > 0
        np.testing.assert_allclose(S, R, rtol=0, atol=1e-10)
        np.testing.assert_allclose(S.conj().T @ S, np.eye(S.shape[0]), atol=1e-10)
