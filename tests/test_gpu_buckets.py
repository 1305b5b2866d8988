"""Bucketed sweep of uneven two-rotor blocks on the device: every block against
its own sweep and against the oracle; the tile walk leaves results bit for bit
unchanged; the sweep autotuner's model; blocks past 1024 channels."""

import numpy as np
import pytest

import pyticc_amd as ticc
from oracle import pipeline
from pyticc_amd import _backend, synthetic
from pyticc_amd.propagation.buckets import (
    WALK_CANDIDATES,
    Bucket,
    padded,
    plan_buckets,
    plan_singletons,
    propagate_buckets,
    tune_sweep,
)
from pyticc_amd.scattering.hamiltonian import build_hamiltonian
from tests.helpers import hamiltonian_problem

pytestmark = pytest.mark.gpu


def _blocks(sizes, Js=(1, 2), n_sector=10):
    """Truncated two-rotor (J, +1) blocks (j_X <= 2, j_Y <= 5) on a short radial grid."""
    hams = {}
    grid = None
    for J in Js:
        w = synthetic.rotor_rotor(jmax_X=2, jmax_Y=5, Jtot=J, parity=1, n_energy=3,
                                  boundaries=(4.0, 4.0 + 2 * 0.009 * n_sector), n_theta_X=8, n_theta_Y=8, n_phi=6)
        if grid is None:
            grid = ticc.prepare_potential(w.system, w.boundaries, w.half_steps, **w.quadrature)
        hams[J] = (build_hamiltonian(w.system, grid), w.energies)
    subs = [hams[Js[i % len(Js)]][0].truncate(n) for i, n in enumerate(sizes)]
    return subs, hams[Js[0]][1], grid


def test_bucketed_sweep_matches_each_block_and_the_oracle(gpu_device):
    sizes = (20, 47, 33, 55, 61)  # J = 2 (111 channels) and J = 1 (57) alternately
    subs, E, grid = _blocks(sizes, Js=(2, 1))
    prop = ticc.Propagation(device="gpu")
    plan = (Bucket(padded(47), (0, 2, 1)), Bucket(padded(61), (3, 4), (256, 1024)))
    Yb = propagate_buckets(subs, E, grid.sectors, prop, plan)
    Ys = propagate_buckets(subs, E, grid.sectors, prop, plan_singletons(sizes))
    for n, sub, a, b in zip(sizes, subs, Yb, Ys):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert a.shape == (E.size, n, n)
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())
        ref = pipeline.propagate(hamiltonian_problem(sub, grid), E)
        np.testing.assert_allclose(a, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_two_lanes_equal_one_lane(gpu_device):
    """Buckets swept concurrently by two host threads on two library lanes and
    streams give the same results, bit for bit, as one after another."""
    sizes = (20, 47, 33, 55, 61, 90, 100)
    subs, E, grid = _blocks(sizes, Js=(2, 1, 2, 1, 2, 2, 2))
    prop = ticc.Propagation(device="gpu")
    plan = (Bucket(padded(47), (0, 2, 1)), Bucket(padded(61), (3, 4)), Bucket(96, (5,)), Bucket(padded(100), (6,)))
    one = propagate_buckets(subs, E, grid.sectors, prop, plan, lanes=1)
    timings = []
    two = propagate_buckets(subs, E, grid.sectors, prop, plan, timings=timings, lanes=2)
    assert len(timings) == len(plan)
    for a, b in zip(one, two):
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    with pytest.raises(ValueError, match="lanes"):
        propagate_buckets(subs, E, grid.sectors, prop, plan, lanes=3)
    with pytest.raises(ValueError):
        _backend.call("ticc_set_lane", 4)


def test_tile_walk_does_not_change_results(gpu_device):
    subs, E, grid = _blocks((90,), Js=(2,))
    prop = ticc.Propagation(device="gpu")
    runs = [propagate_buckets(subs, E, grid.sectors, prop, (Bucket(96, (0,), walk),))[0].cpu().numpy()
            for walk in WALK_CANDIDATES]
    for r in runs[1:]:
        np.testing.assert_array_equal(r, runs[0])


def test_plan_rejects_a_wrong_cover(gpu_device):
    subs, E, grid = _blocks((20, 30))
    prop = ticc.Propagation(device="gpu")
    with pytest.raises(ValueError, match="exactly once"):
        propagate_buckets(subs, E, grid.sectors, prop, (Bucket(32, (0,)),))
    with pytest.raises(ValueError, match="does not fit"):
        propagate_buckets(subs, E, grid.sectors, prop, (Bucket(64, (0, 1)),))


def test_walk_knob_validates(gpu_device):
    with pytest.raises(ValueError):
        _backend.call("ticc_set_sweep_walk", -1, 0)
    _backend.call("ticc_set_sweep_walk", 0, 0)


def test_tuner_model_and_plan(gpu_device):
    cost = tune_sweep((64, 160), batches=(4, 16), device="gpu", walks=WALK_CANDIDATES[:2])
    assert cost.n_pads == (64, 160)
    assert all(a >= 0 for a in cost.a) and all(b > 0 for b in cost.b)
    assert all(w in WALK_CANDIDATES[:2] for w in cost.walks)
    assert len(cost.samples) == 4
    print("tuned", cost.n_pads, cost.a, cost.b, cost.walks)
    sizes = [40, 50, 64, 100, 150, 160]
    plan = plan_buckets(sizes, 8, cost)
    assert sorted(i for b in plan for i in b.blocks) == list(range(len(sizes)))


def test_sweep_past_1024_channels(gpu_device):
    """A truncated J = 8 block of 1100 channels (three panel rows per thread) against the oracle."""
    w = synthetic.rotor_rotor(Jtot=8, parity=1, n_energy=2, boundaries=(4.0, 4.018), n_theta_X=8, n_theta_Y=8,
                              n_phi=6)
    grid = ticc.prepare_potential(w.system, w.boundaries, w.half_steps, **w.quadrature)
    sub = build_hamiltonian(w.system, grid).truncate(1100)
    (Y,) = propagate_buckets([sub], w.energies, grid.sectors, ticc.Propagation(device="gpu"), plan_singletons([1100]))
    ref = pipeline.propagate(hamiltonian_problem(sub, grid), w.energies)
    got = Y.cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11 * np.abs(ref).max())
