"""The mixed-precision form on the device: the FP32 Gauss-Jordan inversion with
FP64-residual Newton refinement (ticc_invert_batched_mixed), the mixed LDMD
sweep against the FP64 oracle (ticc_ldmd_propagate_mixed), and end-to-end S
matrices of the benchmark systems against the oracle fixtures at the precision
form's bar, max |dS| <= 1e-6, with the unitarity defect and the relative change
of the partial cross sections stated."""

from pathlib import Path

import numpy as np
import pytest

import pyticc_amd as ticc
from oracle import ldmd as oldmd
from oracle import matching as omatch
from pyticc_amd import _backend, _device, synthetic
from tests.helpers import random_W_batch, rotor_rotor_N800_reference

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
TOL_MIXED = 1.0e-6


def _stream(dev):
    import torch

    return torch.cuda.current_stream(dev).cuda_stream


def _invert_mixed(gpu_device, A, refine):
    d = _device.to_device(A, gpu_device)
    _backend.call("ticc_invert_batched_mixed", A.shape[0], A.shape[1], refine, d.data_ptr(), _stream(gpu_device))
    return _device.to_host(d)


@pytest.mark.parametrize("N", [1, 31, 100, 800, 1100])
def test_mixed_inverse_converges_with_newton_steps(gpu_device, N):
    """FP32 inversion alone is accurate to the FP32 rounding level; each Newton
    step with an FP64 residual squares the error, so two reach FP64 level."""
    rng = np.random.default_rng(N + 11)
    A = rng.standard_normal((3, N, N)) / np.sqrt(N) + 2.0 * np.eye(N)  # condition number about 3
    ref = np.linalg.inv(A)
    scale = np.abs(ref).max()
    errs = []
    for refine in (0, 1, 2):
        print(f"mixed inverse N={N} refine={refine}", flush=True)
        X = _invert_mixed(gpu_device, A, refine)
        errs.append(float(np.abs(X - ref).max() / scale))
    print(f"N={N}: relative error after 0, 1, 2 Newton steps: {errs}")
    assert 1e-9 < errs[0] < 1e-5  # an FP32 inverse, not an FP64 one
    assert errs[1] < max(1e-13, 10.0 * errs[0] ** 2)  # Newton: the error squares
    assert errs[2] < 1e-13


def test_mixed_inverse_pivots(gpu_device):
    # a zero leading diagonal: without row pivoting the FP32 elimination divides by zero
    A = np.array([[[0.0, 1.0, 2.0], [1.0, 0.0, 3.0], [4.0, 5.0, 0.0]]])
    np.testing.assert_allclose(_invert_mixed(gpu_device, A, 2)[0], np.linalg.inv(A[0]), rtol=1e-14, atol=1e-14)
    rng = np.random.default_rng(5)
    B = rng.standard_normal((2, 200, 200))  # unscaled Gaussian: condition number in the hundreds
    X = _invert_mixed(gpu_device, B, 2)
    np.testing.assert_allclose(X @ B, np.broadcast_to(np.eye(200), B.shape), atol=1e-10)


def _sweep_mixed(gpu_device, Y0, E, mu, hs, Ws, Wm, We, refine):
    N = Y0.shape[-1]
    W = np.empty((2 * hs.size + 1, N, N))
    W[0:-1:2] = Ws
    W[1::2] = Wm
    W[2::2] = We
    dW = _device.to_device(W, gpu_device)
    dY = _device.to_device(Y0, gpu_device)
    E = np.ascontiguousarray(E, dtype=np.float64)
    hs = np.ascontiguousarray(hs, dtype=np.float64)
    base, n2 = dW.data_ptr(), N * N
    _backend.call("ticc_ldmd_propagate_mixed", E.size, N, E.ctypes.data_as(_backend._dp), float(mu), hs.size,
                  hs.ctypes.data_as(_backend._dp), base, base + 8 * n2, base + 16 * n2, 2 * n2, refine, dY.data_ptr(),
                  _stream(gpu_device))
    return _device.to_host(dY)


@pytest.mark.parametrize("N,n_E,n_S", [(3, 2, 4), (64, 3, 6), (150, 5, 5)])
def test_mixed_sweep_matches_oracle(gpu_device, N, n_E, n_S):
    rng = np.random.default_rng(N * 3 + n_S)
    Ws, Wm, We = random_W_batch(rng, n_S, N, scale=20.0)
    # the sweep reads W at consecutive radial points: sector s ends where s + 1 starts
    We[:-1] = Ws[1:]
    hs = rng.uniform(0.01, 0.03, n_S)
    E = rng.uniform(0.5, 2.0, n_E)
    mu = 3.0
    Y0 = oldmd.initialize_inelastic(Ws[0][None] - 2 * mu * E[:, None, None] * np.eye(N))
    ref = oldmd.propagate(Y0, E, mu, hs, Ws, Wm, We)
    scale = np.abs(ref).max()
    got = _sweep_mixed(gpu_device, Y0, E, mu, hs, Ws, Wm, We, 2)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10 * scale)
    np.testing.assert_array_equal(got, np.swapaxes(got, 1, 2))


def _mixed_stats(reset=True):
    inv = np.zeros(1, dtype=np.int64)
    fb = np.zeros(1, dtype=np.int64)
    _backend.call("ticc_mixed_stats", inv.ctypes.data, fb.ctypes.data, int(reset))
    return int(inv[0]), int(fb[0])


def test_mixed_sweep_with_every_matrix_falling_back_is_the_fp64_sweep(gpu_device, monkeypatch):
    """A sector matrix whose refined FP32 inverse would miss FP64 accuracy is
    inverted in FP64 instead; forcing that for every matrix (residual bound 0)
    must reproduce the FP64 sweep bit for bit."""
    from pyticc_amd.propagation.logd import propagate_logD

    rng = np.random.default_rng(17)
    N, n_E, n_S = 96, 3, 6
    Ws, Wm, We = random_W_batch(rng, n_S, N, scale=20.0)
    We[:-1] = Ws[1:]
    hs = rng.uniform(0.01, 0.03, n_S)
    E = rng.uniform(0.5, 2.0, n_E)
    mu = 3.0
    Y0 = oldmd.initialize_inelastic(Ws[0][None] - 2 * mu * E[:, None, None] * np.eye(N))
    ref = _device.to_host(propagate_logD(Y0, E, mu, hs, Ws, Wm, We))
    _mixed_stats(reset=True)
    monkeypatch.setenv("TICC_MIXED_FALLBACK_TAU", "0")
    got = _sweep_mixed(gpu_device, Y0, E, mu, hs, Ws, Wm, We, 2)
    assert _mixed_stats() == (2 * n_S * n_E, 2 * n_S * n_E)
    np.testing.assert_array_equal(got, ref)
    monkeypatch.setenv("TICC_MIXED_FALLBACK_TAU", "1e30")  # never: the refined FP32 inverses throughout
    got = _sweep_mixed(gpu_device, Y0, E, mu, hs, Ws, Wm, We, 2)
    assert _mixed_stats() == (2 * n_S * n_E, 0)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


def _cross_section_change(S, S_ref, res):
    """max over energies and level pairs of |sigma - sigma_ref| / max(sigma_ref) at that energy."""
    from dataclasses import replace

    cs = ticc.partial_cross_sections(replace(res, Smat=tuple(S)))
    cs_ref = ticc.partial_cross_sections(replace(res, Smat=tuple(S_ref)))
    worst = 0.0
    for a, b in zip(cs, cs_ref):
        if b.sigma.size:
            worst = max(worst, float(np.max(np.abs(a.sigma - b.sigma)) / np.max(np.abs(b.sigma))))
    return worst


def test_mixed_solve_atom_rotor_N66_matches_fixture(gpu_device):
    d = np.load(GOLDEN / "smat_atom_rotor_N66.npz")
    w = synthetic.atom_rotor()
    grid = ticc.prepare_potential(w.system, w.boundaries, w.half_steps, **w.quadrature)
    worst = {}
    for refine in (0, 2):
        print(f"mixed N=66 solve, refine={refine}", flush=True)
        res = ticc.solve(w.system, w.energies, grid, ticc.Propagation(device="gpu", precision="mixed",
                                                                       refine_steps=refine))
        idx = [int(e) for e in d["energy_index"]]
        S = [res.Smat[e] for e in idx]
        S_ref = [d[f"S{i}"] for i in range(len(idx))]
        worst[refine] = max(float(np.max(np.abs(a - b))) for a, b in zip(S, S_ref))
        if refine == 2:
            dsig = _cross_section_change(S, S_ref, replace_energies(res, idx))
            defect = max(omatch.unitarity_defect(s) for s in res.Smat if s.size)
    print(f"N = 66 mixed: max|dS| refine 0 {worst[0]:.2e}, refine 2 {worst[2]:.2e}; rel dsigma {dsig:.2e}; "
          f"unitarity {defect:.1e}")
    assert worst[0] > 1e-7  # the FP32 inversions alone are far from FP64 ...
    assert worst[2] <= TOL_MIXED  # ... and two Newton steps meet the bar
    assert dsig <= 1e-5 and defect < 1e-9


def replace_energies(res, idx):
    from dataclasses import replace

    return replace(res, Etot=np.asarray(res.Etot)[idx], Smat=tuple(res.Smat[i] for i in idx))


def test_mixed_solve_rotor_rotor_N800_full_size_matches_fixture(gpu_device):
    d = np.load(GOLDEN / "smat_rotor_rotor_N800.npz")
    w = synthetic.rotor_rotor()
    idx = d["energy_index"]
    grid = ticc.prepare_potential(w.system, w.boundaries, w.half_steps, **w.quadrature)
    _mixed_stats(reset=True)
    res = ticc.solve(w.system, w.energies[idx], grid,
                     ticc.Propagation(device="gpu", memory_mb=4096.0, precision="mixed", refine_steps=2))
    report = []
    for i in range(idx.size):
        S = res.Smat[i]
        assert S.shape[0] == int(d["n_open"][i])
        if S.size == 0:
            continue
        r, c, ref = rotor_rotor_N800_reference(d, i)
        report.append((int(idx[i]), S.shape[0], float(np.max(np.abs(S[r, c] - ref))), omatch.unitarity_defect(S)))
    print("N = 800 mixed (energy, n_open, max|dS|, unitarity defect):", report, "inversions, FP64 fallbacks:",
          _mixed_stats())
    for e, n, err, defect in report:
        assert err <= TOL_MIXED, (e, n, err)
        assert defect < 1e-9, (e, defect)


def test_mixed_sweep_default_two_groups_equal_one_group(gpu_device, monkeypatch):
    """The mixed sweep's default of two energy groups (Np >= 256, 16 or more
    energies) gives the one-group mixed sweep bit for bit, fallbacks included."""
    rng = np.random.default_rng(37)
    N, n_E, n_S = 300, 16, 4
    Ws, Wm, We = random_W_batch(rng, n_S, N, scale=20.0)
    We[:-1] = Ws[1:]
    hs = rng.uniform(0.01, 0.03, n_S)
    E = rng.uniform(0.5, 2.0, n_E)
    mu = 3.0
    Y0 = oldmd.initialize_inelastic(Ws[0][None] - 2 * mu * E[:, None, None] * np.eye(N))
    print(f"mixed default groups: N={N} energies={n_E} sectors={n_S}", flush=True)
    default = _sweep_mixed(gpu_device, Y0, E, mu, hs, Ws, Wm, We, 2)
    monkeypatch.setenv("TICC_SWEEP_GROUPS", "1")
    one = _sweep_mixed(gpu_device, Y0, E, mu, hs, Ws, Wm, We, 2)
    np.testing.assert_array_equal(default, one)
