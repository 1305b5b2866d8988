"""The batching form as one captured HIP graph (CapturedBuckets): every bucket's
sweep enqueued without host work, captured once and replayed.  Each replay must
equal the uncaptured bucketed sweep and every block's own sweep bit for bit,
also with eager device work and synchronisations between replays; the seeded
256-block workload of the batching benchmark runs over its first 2 sectors."""

import numpy as np
import pytest
import torch

import pyticc_amd as ticc
from pyticc_amd import synthetic
from pyticc_amd.propagation.buckets import (
    DEFAULT_WALK,
    Bucket,
    CapturedBuckets,
    SweepCost,
    padded,
    plan_buckets,
    propagate_buckets,
)
from pyticc_amd.propagation.runner import propagate_blocks
from pyticc_amd.scattering.hamiltonian import build_hamiltonian
from tests.test_gpu_buckets import _blocks

pytestmark = pytest.mark.gpu


def _host(Ys):
    return [y.detach().cpu().numpy().copy() for y in Ys]


def test_captured_buckets_replays_equal_the_uncaptured_sweep(gpu_device):
    sizes = (20, 47, 33, 55, 61, 90)
    subs, E, grid = _blocks(sizes, Js=(2, 1, 2, 1, 2, 2))
    prop = ticc.Propagation(device="gpu")
    plan = (Bucket(padded(47), (0, 2, 1)), Bucket(padded(61), (3, 4), (256, 1024)), Bucket(padded(90), (5,)))
    ref = _host(propagate_buckets(subs, E, grid.sectors, prop, plan, lanes=1))
    cap = CapturedBuckets(subs, E, grid.sectors, prop, plan)
    try:
        for replay in range(3):
            print(f"graph replay {replay}", flush=True)
            cap.replay()
            torch.cuda.synchronize()
            got = _host(cap.results())
            for n, a, b in zip(sizes, got, ref):
                assert a.shape == (E.size, n, n)
                np.testing.assert_array_equal(a, b)
            # eager work between replays: a device reduction and a host round trip
            torch.ones(1 << 20, device=gpu_device).sum().item()
    finally:
        cap.close()


def test_captured_256_block_workload_equals_each_blocks_own_sweep(gpu_device):
    specs = synthetic.rotor_rotor_block_specs(seed=3, n_block=256)
    h = 0.009
    works = synthetic.rotor_rotor_blocks(specs, n_energy=32, boundaries=(4.0, 4.0 + 2 * h * 2), half_step=h)
    first = next(iter(works.values()))
    grid = ticc.prepare_potential(first.system, first.boundaries, first.half_steps, **first.quadrature)
    full = {k: build_hamiltonian(w.system, grid) for k, w in works.items()}
    hams = [full[(s.Jtot, s.parity)].truncate(s.n_channel) for s in specs]
    sizes = np.array([s.n_channel for s in specs])
    E = first.energies
    assert len(grid.sectors) == 2 and sizes.min() >= 40 and sizes.max() <= 1200
    # a fixed cost model (latency a per launch sequence, b ~ Np^3 per matrix) makes the plan deterministic
    pads = (64, 1216)
    cost = SweepCost(pads, (2e-3, 6e-3), (1e-9 * 64**3 / 64**2, 1e-9 * 1216**3 / 64**2), (DEFAULT_WALK, DEFAULT_WALK))
    plan = plan_buckets(sizes, E.size, cost)
    assert 1 < len(plan) < len(specs)
    prop = ticc.Propagation(device="gpu", memory_mb=8192.0)
    cap = CapturedBuckets(hams, E, grid.sectors, prop, plan)
    try:
        cap.replay()
        torch.cuda.synchronize()
        got = _host(cap.results())
    finally:
        cap.close()
    worst = 0
    for i, ham in enumerate(hams):
        (own,) = propagate_blocks([ham], E, grid.sectors, prop)
        own = own.cpu().numpy()
        assert got[i].shape == own.shape
        assert np.all(np.isfinite(got[i]))
        worst = max(worst, int(np.count_nonzero(got[i] != own)))
    print(f"256 blocks in {len(plan)} captured buckets: elements differing from the blocks' own sweeps: {worst}")
    assert worst == 0
