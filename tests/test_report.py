"""Text reports (src/pyticc/report.py): the reference's own report checks for the
field-free families (tests/unit/test_report.py:15-54, 77-99, 142-163, 226-275,
300-323, 344-352), re-expressed on this package's result types, plus the K-block
(coupled-states) path of the S-matrix table."""

import numpy as np
import pytest

import pyticc_amd as ticc
from pyticc_amd import report
from pyticc_amd.basis.channel import Channel, ChannelBasis
from pyticc_amd.constants import AU2CM
from pyticc_amd.result import CoupledStatesResult, KBlockResult, ScatteringResult
from pyticc_amd.system import MolInnerState


def _diabatic_basis():
    atom = MolInnerState(j=0)
    states = (MolInnerState(j=1, v=0, Eint=0.001, electronic_state=0), MolInnerState(j=3, v=0, Eint=0.002, electronic_state=0),
              MolInnerState(j=1, v=1, Eint=0.003, electronic_state=1))
    return ChannelBasis(channels=tuple(Channel(atom, s, s.j, 0, s.Eint) for s in states), Jtot=0, system_parity=1)


def _result():
    first = np.arange(9, dtype=np.float64).reshape(3, 3).astype(np.complex128)
    second = first + 10.0 + 1.0j * np.arange(9, dtype=np.float64).reshape(3, 3)
    return ScatteringResult(basis=_diabatic_basis(), Etot=np.array([0.01, 0.02]), Y_propagated=np.zeros((2, 3, 3)),
                            asymptotic_transform=np.eye(3), L=np.array([0.0, 2.0, 1.0]), Smat=(first, second))


def _diatom_diatom_result():
    basis = ChannelBasis(channels=(
        Channel(MolInnerState(v=0, j=0, Eint=0.001), MolInnerState(v=1, j=1, Eint=0.002), 1, 0, 0.003),
        Channel(MolInnerState(v=1, j=2, Eint=0.003), MolInnerState(v=0, j=1, Eint=0.001), 2, 0, 0.004)), Jtot=2, system_parity=1)
    return ScatteringResult(basis=basis, Etot=np.array([0.01]), Y_propagated=np.zeros((1, 2, 2)), asymptotic_transform=np.eye(2),
                            L=np.array([1.0, 2.0]), Smat=(np.array([[1.0, 2.0], [3.0, 4.0]], dtype=np.complex128),))


def test_table_layout():
    out = report._table(("a", "bbb"), [["1", "2"], ["333", "4"]])
    assert out.splitlines() == ["  a  bbb", "---  ---", "  1    2", "333    4"]


def test_channels_omits_block_quantum_numbers_and_includes_electronic_state():
    header = report.channels(_diabatic_basis()).splitlines()[0].split()
    assert header == ["n", "state", "v", "j", "K", "E_int/cm-1"]
    out = report.channels(_diabatic_basis())
    assert "Jtot" not in out and "parity" not in out
    assert out.splitlines()[2].split()[-1] == f"{0.001 * AU2CM:.6f}"


def test_open_closed_uses_cm_inverse():
    out = report.open_closed(_diabatic_basis(), [0.01])
    assert "Etot/cm-1" in out and f"{0.01 * AU2CM:.8f}" in out
    assert out.splitlines()[-1].split()[-2:] == ["3", "0"]
    out = report.open_closed(_diabatic_basis(), [0.0015, 0.0025])
    assert [line.split()[-2:] for line in out.splitlines()[2:]] == [["1", "2"], ["2", "1"]]


def test_k_blocks_reports_membership():
    block = ticc.KBlock(index=0, center_K=1, K_delta=1, K_values=(0, 1, 2), channel_indices=(0, 1, 2),
                        owned_K_values=(1,), owned_channel_indices=(1,))
    out = report.k_blocks([block])
    assert "center_K" in out and "0,1,2" in out and "owned_channels" in out
    assert out.splitlines()[-1].split() == ["0", "1", "0,1,2", "3", "1", "1"]


def test_smatrix_filters_energy_and_initial_and_final_vj_ranges():
    result = _result()
    out = report.smatrix(result, energy_indices=1, state=0, v=0, j=range(1, 2), state_prime=1, v_prime=range(1, 2), j_prime=[1])
    lines = out.splitlines()
    assert lines[0].split() == ["Etot/cm-1", "state", "v", "j", "L", "state'", "v'", "j'", "L'", "Re(S)", "Im(S)"]
    assert len(lines) == 3
    assert f"{result.Etot[1] * AU2CM:.8f}" in lines[-1]
    # S[out = channel 2, in = channel 0] of the second matrix: 16 + 6i
    assert lines[-1].split()[-2:] == ["1.6000000000000000E+01", "6.0000000000000000E+00"]


def test_smatrix_uses_outgoing_row_and_incoming_column():
    out = report.smatrix(_result(), energy_indices=0, state=0, v=0, j=1, state_prime=0, v_prime=0, j_prime=3)
    assert out.splitlines()[-1].split()[-2] == "3.0000000000000000E+00"   # S[1, 0] = 3


def test_smatrix_full_table_order_and_count():
    result = _result()
    lines = report.smatrix(result).splitlines()[2:]
    assert len(lines) == 2 * 3 * 3
    got = [complex(float(x.split()[-2]), float(x.split()[-1])) for x in lines[:9]]
    np.testing.assert_array_equal(got, result.Smat[0].T.ravel())   # incoming outermost


def test_diatom_diatom_smatrix_filters_both_monomers():
    out = report.smatrix(_diatom_diatom_result(), v_X=0, j_X=0, v_Y=1, j_Y=1, j_couple=1, v_X_prime=1, j_X_prime=2,
                         v_Y_prime=0, j_Y_prime=1, j_couple_prime=2)
    assert out.splitlines()[0].split()[1:7] == ["v_X", "j_X", "v_Y", "j_Y", "j_couple", "L"]
    assert len(out.splitlines()) == 3 and "3.0000000000000000E+00" in out


def test_diatom_diatom_smatrix_rejects_single_monomer_filter():
    with pytest.raises(ValueError, match="do not apply to diatom-diatom"):
        report.smatrix(_diatom_diatom_result(), v=0)


def test_smatrix_rejects_unmatched_quantum_numbers_and_bad_selections():
    with pytest.raises(ValueError, match="initial-state selection"):
        report.smatrix(_result(), v=99)
    with pytest.raises(TypeError):
        report.smatrix(_result(), v="0")
    with pytest.raises(ValueError, match="do not apply to atom-diatom"):
        report.smatrix(_result(), v_X=0)


def test_coupled_states_smatrix_reports_one_block():
    exact = _result()
    blk = ticc.KBlock(0, 0, 0, (0,), (0, 1, 2), (0,), (0, 1, 2))
    br = KBlockResult(block=blk, open_channel_indices=exact.open_channel_indices, Y_BF=exact.Y_propagated, Bmat=np.eye(3),
                      L=exact.L, Smat_asymptotic=exact.Smat)
    cs = CoupledStatesResult(basis=exact.basis, Etot=exact.Etot, approx=ticc.Approx.CS, blocks=(br,))
    assert report.smatrix(cs) == report.smatrix(exact)
    two = CoupledStatesResult(basis=exact.basis, Etot=exact.Etot, approx=ticc.Approx.CS, blocks=(br, br))
    with pytest.raises(ValueError, match="block_index is required"):
        report.smatrix(two)
    assert report.smatrix(two, block_index=1) == report.smatrix(exact)
    with pytest.raises(ValueError, match="only valid for coupled-states"):
        report.smatrix(exact, block_index=0)
    odd = KBlockResult(block=blk, open_channel_indices=exact.open_channel_indices, Y_BF=exact.Y_propagated, Bmat=np.eye(3),
                       L=exact.L + 0.3, Smat_asymptotic=exact.Smat)
    with pytest.raises(ValueError, match="must be integral"):
        report.smatrix(CoupledStatesResult(basis=exact.basis, Etot=exact.Etot, approx=ticc.Approx.CS, blocks=(odd,)))


def test_rovib_levels_sorted_by_energy():
    rotor = ticc.rigid_rotor(10.0 / AU2CM, 3)
    lines = report.rovib_levels(rotor).splitlines()
    assert lines[0].split() == ["v", "j", "E_int/cm-1"]
    assert [float(x.split()[-1]) for x in lines[2:]] == pytest.approx([0.0, 20.0, 60.0, 120.0])
