"""Compiled PES through the C-ABI of include/ticc_pes.h (the compiled-PES seam of
src/pyticc/pes/fortran/executor.py:283-397): a small C surface is built with gcc,
loaded, checked against the same formula in NumPy, and used to prepare a PES grid
that must equal the grid of the Python form of the surface."""

import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import pyticc_amd as ticc
from pyticc_amd.constants import AMU2AU, CM2AU
from pyticc_amd.pes.native import load_native_pes

HEADER = Path(__file__).resolve().parents[1] / "include" / "ticc_pes.h"

C_SOURCE = r"""
#include <math.h>
#include "ticc_pes.h"
/* V(R, r, theta) = A exp(-b R) (1 + c cos(theta) + d cos(theta)^2) (1 + e (r - r0)) */
static double v1(double R, double r, double th) {
  const double ct = cos(th);
  return 0.5 * exp(-1.3 * R) * (1.0 + 0.4 * ct + 0.25 * ct * ct) * (1.0 + 0.1 * (r - 1.4));
}
int pes_interaction(double R, const double* coords, int n_coord, int n_grid, double* out) {
  if (n_coord != 2) return 3;
  if (R < 0.0) return 7;
  for (int g = 0; g < n_grid; ++g) out[g] = v1(R, coords[2 * g], coords[2 * g + 1]);
  return 0;
}
#ifdef WITH_MANY
int pes_interaction_many(const double* R, int n_R, const double* coords, int n_coord, int n_grid, double* out) {
  for (int k = 0; k < n_R; ++k) {
    const int s = pes_interaction(R[k], coords, n_coord, n_grid, out + (long)k * n_grid);
    if (s) return s;
  }
  return 0;
}
#endif
int pes_monomer(const double* r, int n, double* out) {
  for (int i = 0; i < n; ++i) out[i] = 0.2 * (r[i] - 1.4) * (r[i] - 1.4);
  return 0;
}
"""


def _py_v(R, coords):
    r, th = coords
    ct = np.cos(th)
    return 0.5 * np.exp(-1.3 * R) * (1.0 + 0.4 * ct + 0.25 * ct * ct) * (1.0 + 0.1 * (r - 1.4))


@pytest.fixture(scope="module", params=[False, True], ids=["single", "batched"])
def native_lib(request, tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    d = tmp_path_factory.mktemp("pes")
    (d / "pes.c").write_text(C_SOURCE)
    so = d / "libpes.so"
    cmd = ["gcc", "-O2", "-shared", "-fPIC", f"-I{HEADER.parent}", str(d / "pes.c"), "-o", str(so), "-lm"]
    if request.param:
        cmd.insert(1, "-DWITH_MANY")
    subprocess.run(cmd, check=True)
    return so, request.param


def test_header_declares_the_loader_entry_points():
    names = set(re.findall(r"^int\s+(pes_\w+)\s*\(", HEADER.read_text(), flags=re.M))
    assert names == {"pes_interaction", "pes_interaction_many", "pes_monomer"}


def test_native_pes_matches_formula(native_lib):
    so, _ = native_lib
    pes = load_native_pes(so, monomer_X="pes_monomer")
    rng = np.random.default_rng(3)
    coords = np.stack([rng.uniform(1.0, 2.0, 57), rng.uniform(0.0, np.pi, 57)])
    np.testing.assert_allclose(pes.interaction(4.2, coords), _py_v(4.2, coords), rtol=1e-14)
    Rs = np.linspace(3.0, 9.0, 11)
    many = pes.interaction_many(Rs, coords)
    assert many.shape == (11, 57)
    np.testing.assert_allclose(many, np.stack([_py_v(R, coords) for R in Rs]), rtol=1e-14)
    threaded = pes._interaction_many_processes(Rs, coords, 4)
    np.testing.assert_array_equal(threaded, many)
    r = np.linspace(1.0, 2.0, 5)
    np.testing.assert_allclose(pes.monomer_X(r), 0.2 * (r - 1.4) ** 2, rtol=1e-15)


def test_native_pes_errors(native_lib):
    so, _ = native_lib
    pes = load_native_pes(so)
    coords = np.ones((2, 3))
    with pytest.raises(RuntimeError, match="status 7"):
        pes.interaction(-1.0, coords)
    with pytest.raises(RuntimeError, match="status 3"):
        pes.interaction(1.0, np.ones((3, 3)))
    with pytest.raises(ValueError, match="does not export"):
        load_native_pes(so, interaction="no_such_routine")
    with pytest.raises(ValueError, match="Cannot load"):
        load_native_pes(so.parent / "missing.so")


def test_prepare_potential_with_native_pes_equals_python_pes(native_lib):
    so, _ = native_lib
    rotor = ticc.rigid_rotor(10.0 * CM2AU, 4)

    def system(pes):
        return ticc.build_ScattSystem(ticc.AtomSpec(), rotor, Jtot=2, system_parity=1, potential=pes, reduced_mass=5.0 * AMU2AU)

    py = ticc.PESWrapper(interaction=_py_v)
    ref = ticc.prepare_potential(system(py), (3.0, 6.0), (0.05,), n_theta=8)
    nat = load_native_pes(so)
    for processes in (1, 3):
        got = ticc.prepare_potential(system(nat), (3.0, 6.0), (0.05,), n_theta=8, processes=processes)
        np.testing.assert_allclose(got.values, ref.values, rtol=1e-14, atol=0)
