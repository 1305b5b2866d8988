"""Bucket planning for the batched sweep of uneven blocks, and basis truncation (CPU)."""

import itertools

import numpy as np
import pytest

from oracle.pipeline import w_points
from pyticc_amd import prepare_potential, synthetic
from pyticc_amd.propagation.buckets import (
    Bucket,
    SweepCost,
    modelled_seconds,
    padded,
    plan_buckets,
    plan_singletons,
)
from pyticc_amd.scattering.hamiltonian import build_hamiltonian
from tests.helpers import hamiltonian_problem


def _cost(a=1e-3, b=1e-9):
    # latency a per column pair, work b * Np^3 per matrix
    n = np.array([32, 256, 512, 1024, 1536])
    return SweepCost(tuple(n), tuple(a * n / 64), tuple(b * n**3.0), tuple((512, 2048) for _ in n))


def _brute_force(sizes, n_energy, cost):
    order = np.argsort(sizes, kind="stable")
    s = np.asarray(sizes)[order]
    n = len(s)
    best = np.inf
    for k in range(n):
        for cuts in itertools.combinations(range(1, n), k):
            edges = (0, *cuts, n)
            t = sum(cost.seconds(padded(s[j - 1]), (j - i) * n_energy) for i, j in zip(edges, edges[1:]))
            best = min(best, t)
    return best


@pytest.mark.parametrize("seed", range(5))
def test_plan_is_optimal_over_contiguous_cuts(seed):
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(40), np.log(1200), 8))).astype(int)
    cost = _cost(a=10 ** rng.uniform(-4, -2), b=10 ** rng.uniform(-12, -10))
    plan = plan_buckets(sizes, 32, cost)
    assert sorted(i for b in plan for i in b.blocks) == list(range(sizes.size))
    for b in plan:
        assert b.n_pad == padded(max(sizes[list(b.blocks)]))
    assert modelled_seconds(plan, 32, cost) == pytest.approx(_brute_force(sizes, 32, cost), rel=1e-12)


def test_plan_extremes():
    sizes = [40, 41, 300, 700, 1100, 1200]
    # pure latency: everything in one launch
    one = plan_buckets(sizes, 32, _cost(a=1.0, b=1e-15))
    assert len(one) == 1 and one[0].n_pad == 1216
    # pure work: padding never pays, equal padded sizes still share a launch
    sep = plan_buckets(sizes, 32, _cost(a=0.0, b=1e-9))
    assert [b.n_pad for b in sep] == [64, 320, 704, 1120, 1216]
    assert sep[0].blocks == (0, 1)


def test_plan_respects_the_memory_bound():
    sizes = [1000] * 6
    plan = plan_buckets(sizes, 32, _cost(a=1.0, b=1e-15), max_batch_bytes=3 * 8 * 64 * 1024 * 1024 + 1)
    assert all(len(b.blocks) <= 2 for b in plan)
    assert sorted(i for b in plan for i in b.blocks) == list(range(6))


def test_plan_rejects_bad_input():
    with pytest.raises(ValueError):
        plan_buckets([], 32, _cost())
    with pytest.raises(ValueError):
        plan_buckets([10, 0], 32, _cost())
    with pytest.raises(ValueError):
        plan_buckets([10], 0, _cost())


def test_singletons_and_cost_model():
    plan = plan_singletons([70, 33])
    assert plan == (Bucket(96, (0,)), Bucket(64, (1,)))
    c = SweepCost((64, 128), (1.0, 2.0), (0.5, 1.5))
    assert c.seconds(96, 2) == pytest.approx(1.5 + 1.0 * 2)
    assert c.seconds(256, 1) == pytest.approx(4.0 + 1.5 * 8)


def test_block_specs_are_seeded_and_in_range():
    specs = synthetic.rotor_rotor_block_specs(seed=3, n_block=64)
    assert specs == synthetic.rotor_rotor_block_specs(seed=3, n_block=64)
    n = np.array([s.n_channel for s in specs])
    assert n.min() >= 40 and n.max() <= 1200
    for s in specs[:6]:
        full = synthetic.rotor_rotor(Jtot=s.Jtot, parity=s.parity, n_energy=1).system.basis.n_channel
        assert s.n_channel <= full


def test_truncated_hamiltonian_is_the_leading_block():
    w = synthetic.rotor_rotor(jmax_X=2, jmax_Y=3, Jtot=2, parity=1, n_energy=2, boundaries=(4.0, 4.2),
                              n_theta_X=6, n_theta_Y=6, n_phi=6)
    grid = prepare_potential(w.system, w.boundaries, w.half_steps, **w.quadrature)
    ham = build_hamiltonian(w.system, grid)
    N = ham.basis.n_channel
    n = N - 7
    sub = ham.truncate(n)
    assert sub.basis.n_channel == n
    # the cut splits at least one helicity K subset
    Ks = [c.K for c in ham.basis]
    assert set(Ks[:n]) & set(Ks[n:])
    rows = np.arange(5)
    W_full = w_points(hamiltonian_problem(ham, grid), rows)
    W_sub = w_points(hamiltonian_problem(sub, grid), rows)
    np.testing.assert_allclose(W_sub, W_full[:, :n, :n], rtol=0, atol=1e-13 * np.abs(W_full).max())
    with pytest.raises(ValueError):
        ham.truncate(N + 1)
    with pytest.raises(ValueError):
        ham.block(range(n))  # CS / NNCC blocks still need whole K subsets


class _StubHam:
    """Just what propagate_buckets validates before any device work."""

    def __init__(self, n):
        self.basis = type("B", (), {"n_channel": n})()


def test_propagate_buckets_validates_the_plan_before_device_work():
    from pyticc_amd.propagation.buckets import propagate_buckets

    hams = [_StubHam(20), _StubHam(40)]
    with pytest.raises(ValueError, match="exactly once"):
        propagate_buckets(hams, [0.1], (), None, (Bucket(32, (0,)),))
    with pytest.raises(ValueError, match="exactly once"):
        propagate_buckets(hams, [0.1], (), None, (Bucket(64, (0, 1)), Bucket(32, (0,))))
    with pytest.raises(ValueError, match="does not fit"):
        propagate_buckets(hams, [0.1], (), None, (Bucket(32, (0, 1)),))
    with pytest.raises(ValueError, match="lanes"):
        propagate_buckets(hams, [0.1], (), None, (Bucket(64, (0, 1)),), lanes=3)


def test_captured_buckets_validate_before_device_work():
    from pyticc_amd import Propagation
    from pyticc_amd.propagation.buckets import CapturedBuckets

    hams = [_StubHam(20), _StubHam(40)]
    plan = (Bucket(64, (0, 1)),)
    with pytest.raises(NotImplementedError, match="FP64 inelastic"):
        CapturedBuckets(hams, [0.1], ("s",), Propagation(device="gpu", precision="mixed"), plan)
    with pytest.raises(NotImplementedError, match="FP64 inelastic"):
        CapturedBuckets(hams, [0.1], ("s",), Propagation(device="gpu", mode="capture"), plan)
    with pytest.raises(RuntimeError, match="no radial windows"):
        CapturedBuckets(hams, [0.1], (), Propagation(device="gpu"), plan)
    with pytest.raises(ValueError, match="exactly once"):
        CapturedBuckets(hams, [0.1], ("s",), Propagation(device="gpu"), (Bucket(64, (1,)),))
