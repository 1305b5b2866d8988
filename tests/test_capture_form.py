"""Real form of complex log derivatives used by the capture mode: Z -> [[Re Z, -Im Z],
[Im Z, Re Z]] carries sums, products, inverses and real diagonal shifts, the
end-of-sector symmetrisation becomes (E + P E^T P)/2 with P = diag(I, -I), and the
whole LDMD sector of the oracle (src/pyticc/propagation/logd.py:168-194) run on the
real form reproduces the complex sector.  CPU only (the device kernels are checked
against the oracle in test_gpu_capture.py)."""

import numpy as np
import torch

from oracle import ldmd
from pyticc_amd.match.capture import block_diag_real, from_real_form, to_real_form


def _phi(Z):
    return to_real_form(torch.from_numpy(Z)).numpy()


def _rand(rng, n, sym=True):
    Z = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return 0.5 * (Z + Z.T) if sym else Z


def test_real_form_is_an_algebra_map():
    rng = np.random.default_rng(1)
    A, B = _rand(rng, 5, False), _rand(rng, 5, False)
    np.testing.assert_allclose(_phi(A) @ _phi(B), _phi(A @ B), atol=1e-12)
    np.testing.assert_allclose(np.linalg.inv(_phi(A)), _phi(np.linalg.inv(A)), atol=1e-10)
    D = np.diag(rng.standard_normal(5))
    np.testing.assert_allclose(_phi(A + D), _phi(A) + block_diag_real(torch.from_numpy(D)).numpy(), atol=0)
    np.testing.assert_allclose(_phi(1j * A), np.block([[0 * D, -np.eye(5)], [np.eye(5), 0 * D]]) @ _phi(A), atol=1e-12)
    Z = from_real_form(torch.from_numpy(_phi(A))).numpy()
    np.testing.assert_array_equal(Z, A)


def test_symmetrisation_rule():
    rng = np.random.default_rng(2)
    Z = _rand(rng, 4, False)
    E = _phi(Z)
    P = np.diag(np.r_[np.ones(4), -np.ones(4)])
    np.testing.assert_allclose(0.5 * (E + P @ E.T @ P), _phi(0.5 * (Z + Z.T)), atol=1e-15)
    # the kernel's element rule: the transposed element enters with -1 across the halves
    i, j = np.indices(E.shape)
    flip = (i < 4) != (j < 4)
    np.testing.assert_allclose(0.5 * (E + np.where(flip, -E.T, E.T)), _phi(0.5 * (Z + Z.T)), atol=1e-15)


def _sector_real_form(Yr, h, ref, Qs, Qm, Qe, n):
    """The oracle sector on the real form, with the real-form symmetrisation."""
    two = lambda M: block_diag_real(torch.from_numpy(M)).numpy()  # noqa: E731
    y1, y2 = ldmd.reference_values(h, ref)
    Y1, Y2 = np.r_[y1, y1], np.r_[y2, y2]
    A1 = Yr + two(Qs) + np.diag(Y1)
    X1 = np.linalg.solve(A1, np.diag(Y2))
    Ym = two(Qm) + np.diag(Y1) - Y2[:, None] * X1
    A2 = Ym + two(Qm) + np.diag(Y1)
    X2 = np.linalg.solve(A2, np.diag(Y2))
    Ye = two(Qe) + np.diag(Y1) - Y2[:, None] * X2
    i, j = np.indices(Ye.shape)
    return 0.5 * (Ye + np.where((i < n) != (j < n), -Ye.T, Ye.T))


def test_oracle_sector_on_real_form_equals_complex_sector():
    rng = np.random.default_rng(3)
    n, h = 4, 0.05
    W = rng.standard_normal((n, n))
    W = 0.5 * (W + W.T) - np.diag([30.0, 5.0, -10.0, -40.0])
    Y0 = ldmd.initialize_capture(W)
    ref, Qs, Qm, Qe = ldmd.correction_matrices(h, W, W + 0.1, W + 0.2)
    got = _sector_real_form(_phi(Y0), h, ref, Qs, Qm, Qe, n)
    want = ldmd.sector_step(Y0, h, ref, Qs, Qm, Qe)
    np.testing.assert_allclose(from_real_form(torch.from_numpy(got)).numpy(), want, rtol=1e-12, atol=1e-12)
