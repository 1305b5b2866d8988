"""Learned PES surrogate on the device: the BF16 MFMA GEMM epilogues against NumPy,
forward / loss / gradients of the training kernels against the CPU oracle network
(oracle/mlp.py, same BF16 rounding points), a short training run on samples of the
two-rotor PES, and the surrogate used as the PES callback of prepare_potential
(exact PES fallback above 0.05 Eh) against the exact grid."""

import numpy as np
import pytest
import torch

import pyticc_amd as ticc
from oracle import mlp as omlp
from pyticc_amd import _backend, _device, synthetic
from pyticc_amd.surrogate import PESSurrogate

pytestmark = pytest.mark.gpu


def _bf(x):
    return torch.as_tensor(omlp.bf16(x), dtype=torch.float32).to(torch.bfloat16)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_mlp_gemm_epilogues(gpu_device, mode):
    rng = np.random.default_rng(mode)
    M, N, K = 128, 64, 48
    A = omlp.bf16(rng.standard_normal((M, K)).astype(np.float32))
    B = omlp.bf16(rng.standard_normal((N, K)).astype(np.float32))
    bias = rng.standard_normal(N).astype(np.float32)
    Zin = rng.standard_normal((M, N)).astype(np.float32)
    dA, dB = _bf(A).to(gpu_device), _bf(B).to(gpu_device)
    db = torch.as_tensor(bias, device=gpu_device)
    dZ = torch.as_tensor(Zin, device=gpu_device)
    Cf = torch.empty((M, N), dtype=torch.float32, device=gpu_device)
    Cb = torch.empty((M, N), dtype=torch.bfloat16, device=gpu_device)
    CbT = torch.empty((N, M), dtype=torch.bfloat16, device=gpu_device)
    _backend.call("ticc_mlp_gemm", mode, M, N, K, _device.ptr(dA), _device.ptr(dB), _device.ptr(db), _device.ptr(dZ),
                  _device.ptr(Cf), _device.ptr(Cb), _device.ptr(CbT), _device.stream_handle(gpu_device))
    C = A @ B.T
    got = Cf.cpu().numpy().astype(np.float64)
    if mode == 0:
        want = C
    elif mode == 2:
        want = C + bias
    elif mode == 1:
        want = C + bias
        act = omlp.silu(want)
        np.testing.assert_allclose(Cb.float().cpu().numpy(), omlp.bf16(act.astype(np.float32)), rtol=1e-2, atol=1e-2)
        np.testing.assert_array_equal(CbT.float().cpu().numpy(), Cb.float().cpu().numpy().T)
    else:
        want = C * omlp.dsilu(Zin.astype(np.float64))
        np.testing.assert_array_equal(CbT.float().cpu().numpy(), Cb.float().cpu().numpy().T)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5 * np.abs(want).max())


def _samples(n, seed=0, cutoff=None):
    """n geometries (1/R uniform over 4-40 bohr, isotropic angles) and the exact two-rotor
    PES there; with `cutoff`, only points with |V| <= cutoff."""
    pes = synthetic.RotorRotorPES.seeded(2)
    rng = np.random.default_rng(seed)
    out, V = [], []
    while sum(len(v) for v in V) < n:
        R = 1.0 / rng.uniform(1.0 / 40.0, 1.0 / 4.0)
        m = 256
        ang = np.stack([np.full(m, 1.0), np.full(m, 1.0), np.arccos(rng.uniform(-1, 1, m)), np.arccos(rng.uniform(-1, 1, m)),
                        rng.uniform(0, 2 * np.pi, m)])
        v = pes.interaction(R, ang)
        keep = np.abs(v) <= cutoff if cutoff is not None else np.ones(m, dtype=bool)
        out.append(np.stack([np.full(keep.sum(), R), ang[2, keep], ang[3, keep], ang[4, keep]], axis=1))
        V.append(v[keep])
    return np.concatenate(out)[:n], np.concatenate(V)[:n]


def test_training_kernels_match_oracle_gradients(gpu_device):
    X4, V = _samples(300, seed=1, cutoff=0.05)
    net = PESSurrogate(seed=3, device=gpu_device.index)
    loss, gW, gb = net.loss_and_grads(X4, V)
    Ws, bs = net.parameters_host()
    F = omlp.features(X4, net.r_scale)
    ref_loss, rW, rb = omlp.loss_and_grads(F, V / net.v_scale, Ws, bs)
    np.testing.assert_allclose(loss, ref_loss, rtol=2e-3)
    for g, r in zip(gW + gb, rW + rb):
        assert g.shape == r.shape
        np.testing.assert_allclose(g, r, rtol=0, atol=2e-2 * np.abs(r).max() + 1e-12)
    # forward values against the oracle: equal up to FP32 accumulation order, except
    # where an activation rounds to the neighbouring BF16 value (expf vs exp)
    y = net.predict(X4)
    yr = omlp.forward(F, Ws, bs)[0] * net.v_scale
    d = np.abs(y - yr) / np.abs(yr).max()
    assert np.median(d) < 1e-5 and d.max() < 2e-2, (np.median(d), d.max())


def test_training_reduces_loss(gpu_device):
    X4, V = _samples(40000, seed=2)
    net = PESSurrogate(seed=0, device=gpu_device.index)
    hist = net.fit(X4, V, epochs=10, batch=1024, lr=2e-3)
    assert hist[-1] < 0.6 * hist[0], hist
    Xt, Vt = _samples(4000, seed=9, cutoff=0.05)
    rms = float(np.sqrt(np.mean((net.predict(Xt) - Vt) ** 2)))
    vrms = float(np.sqrt(np.mean(Vt**2)))
    print(f"surrogate: loss {hist[0]:.3e} -> {hist[-1]:.3e}, test RMS {rms:.3e} Eh (V RMS {vrms:.3e}) over |V| <= 0.05 Eh")
    assert rms < 0.9 * vrms


def test_surrogate_as_pes_callback(gpu_device):
    X4, V = _samples(40000, seed=4)
    net = PESSurrogate(seed=1, device=gpu_device.index)
    net.fit(X4, V, epochs=6, batch=1024, lr=2e-3)
    w = synthetic.rotor_rotor(jmax_X=1, jmax_Y=2, Jtot=1, n_energy=2, boundaries=(4.0, 6.0))
    exact = w.system.potential
    grid = ticc.prepare_potential(w.system, w.boundaries, w.half_steps, n_theta_X=6, n_theta_Y=6, n_phi=5)
    from dataclasses import replace
    surrogate = net.pes(exact)
    sys_s = replace(w.system, potential=surrogate)
    grid_s = ticc.prepare_potential(sys_s, w.boundaries, w.half_steps, n_theta_X=6, n_theta_Y=6, n_phi=5)
    Ve, Vs = grid.values, grid_s.values
    wall = np.abs(Ve) > 0.05
    np.testing.assert_allclose(Vs[wall], Ve[wall], rtol=1e-13)   # exact PES on the repulsive wall
    err = np.abs(Vs - Ve)[~wall]
    print(f"surrogate grid: max |dV| {err.max():.3e} Eh off the wall; exact PES at {net.fallback_points} of "
          f"{net.fallback_points + net.surrogate_points} points (wall: {int(wall.sum())})")
    assert err.max() < 0.05


# -- the angular-basis form -------------------------------------------------------------

from pyticc_amd.surrogate import PESBasisSurrogate, angular_basis, angular_terms  # noqa: E402


from pyticc_amd.surrogate import grid_angular_coordinates as _grid_coords  # noqa: E402


def _device_host(t):
    return t.detach().cpu().numpy().astype(np.float64)


def test_basis_head_gradients_match_oracle(gpu_device):
    X4, V = _samples(300, seed=5, cutoff=0.05)
    terms = angular_terms(3)
    net = PESBasisSurrogate(terms, seed=2, device=gpu_device.index)
    C = np.stack([np.ones(len(V)), np.ones(len(V)), X4[:, 1], X4[:, 2], X4[:, 3]])
    G = angular_basis(C, terms)
    loss, gW, gb = net.loss_and_grads_basis(X4[:, 0], G, V)
    Ws, bs = net.parameters_host()
    F = omlp.radial_features(X4[:, 0], net.r_scale)
    T = V / (net.v_scale * net.envelope(X4[:, 0]))
    ref_loss, rW, rb = omlp.basis_loss_and_grads(F, G, T, Ws, bs, split=True)
    np.testing.assert_allclose(loss, ref_loss, rtol=2e-3)
    for g, r in zip(gW + gb, rW + rb):
        assert g.shape == r.shape
        np.testing.assert_allclose(g, r, rtol=0, atol=2e-2 * np.abs(r).max() + 1e-12)
    # the split-BF16 forward: coefficients to about 16 bits, against the oracle's split forward
    U = _device_host(net.coefficients(X4[:, 0]))[:, : len(terms)]
    Uref = omlp.forward(F, Ws, bs, split=True)[1][1][-1]
    np.testing.assert_allclose(U, Uref, rtol=0, atol=2e-5 * np.abs(Uref).max())


@pytest.mark.timeout(400)
def test_basis_surrogate_reproduces_the_pes_and_the_smatrix(gpu_device):
    """Trained on a quarter of the points of a two-rotor PotentialGrid (|V| <= 0.05 Eh;
    400 epochs, about 88k Adam steps), the expansion's held-out error off the wall is
    below 0.5 % of the PES's RMS (about 0.06 % measured), and a solve on the surrogate
    grid (exact PES on the wall and 2 bohr beyond it, about 8 % of the points) stays
    within 1e-3 of the exact grid's S (about 3.5e-4 measured).  The forward products
    run in split BF16 (about 16 bits); plain BF16 products reach about 0.14 % held-out
    with the same schedule."""
    from dataclasses import replace

    w = synthetic.rotor_rotor(jmax_X=2, jmax_Y=3, Jtot=2, n_energy=4)
    quad = {"n_theta_X": 8, "n_theta_Y": 8, "n_phi": 7}
    grid = ticc.prepare_potential(w.system, w.boundaries, w.half_steps, **quad)
    V = grid.values.reshape(grid.radial_points.size, -1)
    terms = angular_terms(6)
    G = angular_basis(_grid_coords(grid), terms)
    rng = np.random.default_rng(7)
    pick = rng.choice(V.size, size=V.size // 4, replace=False)
    ri, gi = np.divmod(pick, V.shape[1])
    net = PESBasisSurrogate(terms, seed=1, device=gpu_device.index)
    hist = net.fit(grid.radial_points[ri], gi, G, V.reshape(-1)[pick], epochs=400, batch=2048, lr=2e-3, lr_decay=0.989)
    pred = net.predict_basis(grid.radial_points, G)
    held = np.ones(V.size, dtype=bool)
    held[pick] = False
    off = held & (np.abs(V.reshape(-1)) <= 0.05)
    rms = float(np.sqrt(np.mean((pred.reshape(-1) - V.reshape(-1))[off] ** 2)))
    vrms = float(np.sqrt(np.mean(V.reshape(-1)[off] ** 2)))
    grid_s = ticc.prepare_potential(replace(w.system, potential=net.pes(w.system.potential)), w.boundaries,
                                    w.half_steps, **quad)
    prop = ticc.Propagation(device="gpu")
    ref = ticc.solve(w.system, w.energies, grid, prop)
    sur = ticc.solve(w.system, w.energies, grid_s, prop)
    dS = max(float(np.abs(a - b).max()) for a, b in zip(ref.Smat, sur.Smat) if a.size)
    frac = net.fallback_points / (net.fallback_points + net.surrogate_points)
    print(f"basis surrogate: loss {hist[0]:.2e} -> {hist[-1]:.2e}; held-out RMS {rms:.2e} Eh of V RMS {vrms:.2e} "
          f"({rms / vrms:.2%}); exact PES at {frac:.1%} of the grid; N={w.system.basis.n_channel} solve max|dS| {dS:.2e}")
    assert rms <= 0.005 * vrms
    assert dS <= 1e-3
