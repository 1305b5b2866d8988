"""CS / NNCC K-block solves on the device (solver.py:154-177, runner.py:326-374,
finalize.py:168-225): every block against the oracle solve of the same block, and
the limits where an approximation is exact CC (the reference's
tests/unit/test_approximation_limits.py: Jtot = 0; K_cut = 0; one NNCC window
covering every K), where the block result must equal the exact one."""

import numpy as np
import pytest

import pyticc_amd as ticc
from oracle import pipeline
from pyticc_amd import synthetic
from pyticc_amd.basis.monomer import rigid_rotor
from pyticc_amd.constants import AMU2AU, CM2AU
from pyticc_amd.scattering.solver import build_k_blocks
from tests.helpers import hamiltonian_problem, oracle_problem

pytestmark = pytest.mark.gpu

TOL_S = 1.0e-10


def _solve(w, grid):
    return ticc.solve(w.system, w.energies, grid, ticc.Propagation(device="gpu"))


def _assert_single_block_equals_exact(exact, approx):
    assert isinstance(exact, ticc.ScatteringResult) and isinstance(approx, ticc.CoupledStatesResult)
    assert approx.basis.channels == exact.basis.channels
    (block,) = approx.blocks
    assert block.block.channel_indices == tuple(range(exact.basis.n_channel))
    np.testing.assert_allclose(block.Y_BF, exact.Y_propagated, rtol=0, atol=1e-13 * np.abs(exact.Y_propagated).max())
    np.testing.assert_allclose(block.Bmat, exact.asymptotic_transform, atol=1e-13)
    np.testing.assert_allclose(block.L, exact.L, atol=1e-13)
    for S_block, S_exact, idx_block, idx_exact in zip(block.Smat_asymptotic, exact.Smat, block.open_channel_indices,
                                                      exact.open_channel_indices):
        np.testing.assert_array_equal(idx_block, idx_exact)
        np.testing.assert_allclose(S_block, S_exact, atol=1e-13)


def _rotor_pair(Jtot, approx, K_cut=None, K_delta=1):
    """Two small rigid rotors, 3 energies, 120 sectors."""
    pes = synthetic.RotorRotorPES.seeded(5).wrapper()
    system = ticc.build_ScattSystem(rigid_rotor(10.0 * CM2AU, 2), rigid_rotor(7.3 * CM2AU, 3), Jtot=Jtot,
                                    system_parity=1, approx=approx, K_delta=K_delta,
                                    channel=ticc.ChannelSpec(K_cut=K_cut), potential=pes, reduced_mass=10.0 * AMU2AU)
    grid = ticc.prepare_potential(system, (4.0, 10.0), (0.025,), n_theta_X=8, n_theta_Y=8, n_phi=6)
    return system, grid


def _rotor_pair_solve(Jtot, approx, **kw):
    system, grid = _rotor_pair(Jtot, approx, **kw)
    return ticc.solve(system, np.array([55.0, 155.0, 405.0]) * CM2AU, grid, ticc.Propagation(device="gpu"))


def test_J_zero_exact_cs_and_nncc_are_identical(gpu_device):
    exact = _rotor_pair_solve(0, ticc.Approx.EXACT)
    _assert_single_block_equals_exact(exact, _rotor_pair_solve(0, ticc.Approx.CS))
    _assert_single_block_equals_exact(exact, _rotor_pair_solve(0, ticc.Approx.NNCC))


def test_K_cut_zero_exact_and_cs_are_identical(gpu_device):
    exact = _rotor_pair_solve(2, ticc.Approx.EXACT, K_cut=0)
    _assert_single_block_equals_exact(exact, _rotor_pair_solve(2, ticc.Approx.CS, K_cut=0))


def test_nncc_covering_every_K_is_identical_to_exact(gpu_device):
    exact = _rotor_pair_solve(2, ticc.Approx.EXACT)
    assert {ch.K for ch in exact.basis} == {0, 1, 2}
    _assert_single_block_equals_exact(exact, _rotor_pair_solve(2, ticc.Approx.NNCC, K_delta=1))


@pytest.mark.parametrize("approx,jmax,Jtot", [(ticc.Approx.CS, 10, 10), (ticc.Approx.NNCC, 8, 6)])
def test_blocks_match_oracle(gpu_device, approx, jmax, Jtot):
    w = synthetic.atom_rotor(jmax=jmax, Jtot=Jtot, n_energy=4, boundaries=(3.0, 20.0), approx=approx)
    grid, ham, _ = oracle_problem(w)
    res = _solve(w, grid)
    assert isinstance(res, ticc.CoupledStatesResult) and res.approx is approx
    blocks = build_k_blocks(ham)
    assert tuple(b.block for b in res.blocks) == blocks
    assert sorted(i for b in blocks for i in b.owned_channel_indices) == list(range(ham.basis.n_channel))
    worst = 0.0
    for br in res.blocks:
        idx = np.asarray(br.block.channel_indices)
        Y, B, L, S = pipeline.solve(hamiltonian_problem(ham.block(idx), grid), w.energies)
        np.testing.assert_allclose(br.Y_BF, Y, rtol=0, atol=1e-10 * np.abs(Y).max())
        np.testing.assert_allclose(br.Bmat, B, atol=1e-13)
        np.testing.assert_allclose(br.L, L, atol=1e-12)
        for e, E in enumerate(w.energies):
            np.testing.assert_array_equal(br.open_channel_indices[e], idx[ham.E_int[idx] < E])
            assert br.Smat_asymptotic[e].shape == S[e].shape
            if S[e].size:
                worst = max(worst, float(np.max(np.abs(br.Smat_asymptotic[e] - S[e]))))
                SBF = br.Smat_BF[e]
                np.testing.assert_allclose(SBF @ SBF.conj().T, np.eye(SBF.shape[0]), atol=1e-9)
    print(f"{approx.value}: {len(blocks)} blocks, max |dS| = {worst:.3e}")
    assert worst <= TOL_S, worst


def test_batched_block_sweep_matches_per_block_sweeps(gpu_device):
    """Blocks of 7, 12 and 20 channels in one padded sweep (ticc_ldmd_propagate_blocks)
    against each block swept alone."""
    from pyticc_amd import _backend, _device
    from pyticc_amd.propagation.logd import propagate_logD
    from tests.helpers import random_W_batch
    import torch

    rng = np.random.default_rng(17)
    sizes, nE, nS, Nmax = (7, 12, 20), 3, 9, 20
    E = np.array([0.4, 0.9, 1.6])
    mu = 2.0
    hs = np.full(nS, 0.02)
    pad = 2 * mu * E.max() + 1.0e4
    Wall = np.zeros((len(sizes), 2 * nS + 1, Nmax, Nmax))
    want = []
    for b, n in enumerate(sizes):
        Ws, Wm, We = random_W_batch(rng, nS, n, scale=15.0)
        pts = np.empty((2 * nS + 1, n, n))  # consecutive radial points: sector s starts where s - 1 ends
        pts[0], pts[1::2], pts[2::2] = Ws[0], Wm, We
        Ws = pts[0:-1:2].copy()
        Wall[b, :, :n, :n] = pts
        Wall[b, :, np.arange(n, Nmax), np.arange(n, Nmax)] = pad
        W0 = Ws[0][None] - 2 * mu * E[:, None, None] * np.eye(n)
        Y0 = np.stack([np.diag(np.sqrt(np.abs(np.diag(w)))) for w in W0])
        want.append(_device.to_host(propagate_logD(Y0, E, mu, hs, Ws, Wm, We)))
    Wd = _device.to_device(Wall, gpu_device)
    Y = torch.zeros((len(sizes) * nE, Nmax, Nmax), dtype=torch.float64, device=gpu_device)
    for b, n in enumerate(sizes):
        d = Wall[b, 0].diagonal()[None, :] - 2 * mu * E[:, None]
        for e in range(nE):
            Y[b * nE + e] = torch.diag(torch.as_tensor(np.sqrt(np.abs(d[e])), device=gpu_device))
    n2 = Nmax * Nmax
    base = _device.ptr(Wd)
    _backend.call("ticc_ldmd_propagate_blocks", len(sizes), nE, Nmax, E.ctypes.data_as(_backend._dp), mu, nS,
                  hs.ctypes.data_as(_backend._dp), base, base + 8 * n2, base + 16 * n2, 2 * n2, (2 * nS + 1) * n2,
                  _device.ptr(Y), _device.stream_handle(gpu_device))
    got = _device.to_host(Y)
    for b, n in enumerate(sizes):
        blk = got[b * nE : (b + 1) * nE]
        np.testing.assert_allclose(blk[:, :n, :n], want[b], rtol=0, atol=1e-12 * np.abs(want[b]).max())
        if n < Nmax:  # the padding channels stay decoupled
            assert np.abs(blk[:, :n, n:]).max() == 0.0 and np.abs(blk[:, n:, :n]).max() == 0.0


def test_batched_blocks_time(gpu_device):
    """The CS blocks of the N = 66 atom + rotor system: one batched sweep against one sweep per block."""
    import time
    from pyticc_amd.propagation.runner import propagate, propagate_blocks

    w = synthetic.atom_rotor(approx=ticc.Approx.CS, n_energy=20)
    grid, ham, _ = oracle_problem(w)
    subs = [ham.block(b.channel_indices) for b in build_k_blocks(ham)]
    prop = ticc.Propagation(device="gpu")
    propagate_blocks(subs, w.energies, grid.sectors[:40], prop)
    torch_sync = __import__("torch").cuda.synchronize
    t = time.perf_counter()
    Yb = propagate_blocks(subs, w.energies, grid.sectors, prop)
    torch_sync()
    t_batched = time.perf_counter() - t
    t = time.perf_counter()
    Ys = [propagate(s, w.energies, grid.sectors, prop) for s in subs]
    torch_sync()
    t_seq = time.perf_counter() - t
    print(f"CS, {len(subs)} blocks x 20 energies x {len(grid.sectors)} sectors: batched {t_batched:.2f} s, "
          f"one sweep per block {t_seq:.2f} s")
    for a, b in zip(Yb, Ys):
        ref = b.cpu().numpy()
        np.testing.assert_allclose(a.cpu().numpy(), ref, rtol=0, atol=1e-11 * np.abs(ref).max())
