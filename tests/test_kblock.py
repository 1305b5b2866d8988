"""CS / NNCC propagation blocks (src/pyticc/basis/kblock.py:53-155) and the block
Hamiltonian: known answers of the reference's own block tests
(tests/unit/test_kblock.py: diatom j <= 4, Jtot = 4) and the restriction invariant
W_block(R) = W_full(R)[block, block] of the contracted oracle W."""

import numpy as np
import pytest

import pyticc_amd as ticc
from oracle import pipeline
from pyticc_amd import synthetic
from pyticc_amd.basis.channel import build_ChannelBasis
from pyticc_amd.basis.monomer import AtomSpec, DiatomSpec
from pyticc_amd.scattering.solver import build_k_blocks
from pyticc_amd.system import ChannelSpec, ScattSystem
from tests.helpers import hamiltonian_problem, oracle_problem


def _channels(Jtot=4, jmax=4):
    system = ScattSystem(AtomSpec(), DiatomSpec(Eint=np.zeros((1, jmax + 1))), Jtot=Jtot, system_parity=1)
    return build_ChannelBasis(system, ChannelSpec())


def test_cs_one_block_per_K_owning_itself():
    channels = _channels()
    blocks = ticc.build_cs_blocks(channels)
    assert [b.K_values for b in blocks] == [(0,), (1,), (2,), (3,), (4,)]
    assert [b.owned_K_values for b in blocks] == [(0,), (1,), (2,), (3,), (4,)]
    assert sorted(i for b in blocks for i in b.owned_channel_indices) == list(range(len(channels)))
    for b in blocks:
        assert all(channels[i].K == b.center_K for i in b.channel_indices)


def test_nncc_delta_one_overlapping_windows():
    channels = _channels()
    blocks = ticc.build_nncc_blocks(channels, K_delta=1)
    assert [b.K_values for b in blocks] == [(0, 1, 2), (1, 2, 3), (2, 3, 4)]
    assert [b.owned_K_values for b in blocks] == [(0, 1), (2,), (3, 4)]
    assert sorted(i for b in blocks for i in b.owned_channel_indices) == list(range(len(channels)))
    assert all(set(b.owned_channel_indices) <= set(b.channel_indices) for b in blocks)
    for b in blocks:
        assert b.channel_indices == tuple(i for i, ch in enumerate(channels) if ch.K in b.K_values)


def test_nncc_wide_window_is_one_block():
    channels = _channels()
    (block,) = ticc.build_nncc_blocks(channels, K_delta=2)
    assert block.K_values == (0, 1, 2, 3, 4) and block.owned_K_values == block.K_values
    assert block.channel_indices == tuple(range(len(channels)))
    (block,) = ticc.build_nncc_blocks(_channels(Jtot=2), K_delta=1)
    assert block.center_K == 1 and block.channel_indices == tuple(range(len(_channels(Jtot=2))))


@pytest.mark.parametrize("K_delta", [0, -1, 1.5, True])
def test_nncc_rejects_bad_delta(K_delta):
    with pytest.raises(ValueError):
        ticc.build_nncc_blocks(_channels(), K_delta=K_delta)


def test_empty_channel_list_has_no_blocks():
    assert ticc.build_cs_blocks(()) == ()
    assert ticc.build_nncc_blocks((), 1) == ()


@pytest.mark.parametrize("approx", [ticc.Approx.CS, ticc.Approx.NNCC])
def test_block_hamiltonian_is_the_principal_submatrix(approx):
    w = synthetic.atom_rotor(jmax=6, Jtot=4, boundaries=(3.0, 4.0), approx=approx)
    grid, ham, prob = oracle_problem(w)
    blocks = build_k_blocks(ham)
    assert len(blocks) == (5 if approx is ticc.Approx.CS else 3)
    rows = np.array([0, 7, 40])
    W_full = pipeline.w_points(prob, rows)
    for block in blocks:
        idx = np.asarray(block.channel_indices)
        sub = ham.block(block.channel_indices)
        np.testing.assert_array_equal(sub.U, ham.U[np.ix_(idx, idx)])
        np.testing.assert_array_equal(sub.E_int, ham.E_int[idx])
        if approx is ticc.Approx.CS:  # one K: no Coriolis coupling inside the block
            np.testing.assert_array_equal(sub.U, np.diag(np.diag(sub.U)))
        W_block = pipeline.w_points(hamiltonian_problem(sub, grid), rows)
        np.testing.assert_allclose(W_block, W_full[:, idx][:, :, idx], rtol=0, atol=1e-12 * np.abs(W_full).max())


def test_block_restriction_rejects_partial_K_subsets():
    w = synthetic.atom_rotor(jmax=4, Jtot=2, boundaries=(3.0, 4.0))
    _, ham, _ = oracle_problem(w)
    K1 = [i for i, ch in enumerate(ham.basis) if ch.K == 1]
    with pytest.raises(ValueError):
        ham.block(K1[:-1])
