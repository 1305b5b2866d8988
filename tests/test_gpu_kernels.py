"""GPU checks of the building blocks: MFMA GEMM, batched GJ inversion, Bessel functions, contraction."""

import numpy as np
import pytest
import torch
from scipy import special

from oracle import bessel as obessel
from oracle import contraction as ocontract
from pyticc_amd import _backend, _device
from pyticc_amd.match.bessel import k_logd_batch, riccati_bessel_batch
from pyticc_amd.matrix.interaction import contract_w_device, prepare_atom_diatom, prepare_diatom_diatom

pytestmark = pytest.mark.gpu


def _stream(dev):
    return _device.stream_handle(dev)


@pytest.mark.parametrize("m,n,k,batch", [(32, 32, 4, 1), (64, 96, 32, 3), (800, 800, 800, 2)])
def test_dgemm_batched_matches_numpy(gpu_device, m, n, k, batch):
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((batch, m, k))
    B = rng.standard_normal((batch, k, n))
    C = rng.standard_normal((batch, m, n))
    dA, dB, dC = (_device.to_device(x, gpu_device) for x in (A, B, C))
    _backend.call("ticc_dgemm_batched", m, n, k, 1.5, dA.data_ptr(), m * k, dB.data_ptr(), k * n, -0.5, dC.data_ptr(), m * n, batch, _stream(gpu_device))
    expect = 1.5 * A @ B - 0.5 * C
    np.testing.assert_allclose(_device.to_host(dC), expect, rtol=1e-12, atol=1e-12 * np.sqrt(k))


@pytest.mark.parametrize("N,batch", [(1, 1), (2, 3), (31, 2), (33, 4), (66, 5), (100, 3), (130, 2), (200, 3), (800, 2),
                                     (1000, 1), (1024, 2), (1025, 2), (1200, 3), (1536, 1)])
def test_invert_batched_matches_numpy(gpu_device, N, batch):
    rng = np.random.default_rng(N)
    A = rng.standard_normal((batch, N, N)) + np.eye(N) * 0.1
    d = _device.to_device(A, gpu_device)
    _backend.call("ticc_invert_batched", batch, N, d.data_ptr(), _stream(gpu_device))
    inv = _device.to_host(d)
    ref = np.linalg.inv(A)
    cond = np.linalg.cond(A).max()
    np.testing.assert_allclose(inv, ref, rtol=0, atol=1e-13 * cond * np.abs(ref).max())
    np.testing.assert_allclose(inv @ A, np.broadcast_to(np.eye(N), A.shape), atol=1e-12 * cond)


def test_invert_above_maximum_size_is_rejected(gpu_device):
    with pytest.raises(NotImplementedError, match="exceeds"):
        _backend.call("ticc_invert_batched", 1, 1537, None, None)


@pytest.mark.parametrize("N", [40, 64, 800])
def test_invert_singular_zero_column_gives_nonfinite_not_a_fault(gpu_device, N):
    """A column of exact zeros leaves every pivot candidate at |a| = 0: the
    inversion must still pick a pivot (value 0) and return non-finite values, as
    LAPACK reports a singular matrix, while the regular matrix of the same batch
    is inverted exactly as alone."""
    rng = np.random.default_rng(N + 1)
    A = rng.standard_normal((3, N, N)) + np.eye(N) * 0.1
    A[0] = 0.0
    A[1][:, N // 3] = 0.0
    d = _device.to_device(A, gpu_device)
    _backend.call("ticc_invert_batched", 3, N, d.data_ptr(), _stream(gpu_device))
    inv = _device.to_host(d)
    assert not np.all(np.isfinite(inv[0])) and not np.all(np.isfinite(inv[1]))
    alone = _device.to_device(A[2:], gpu_device)
    _backend.call("ticc_invert_batched", 1, N, alone.data_ptr(), _stream(gpu_device))
    np.testing.assert_array_equal(inv[2], _device.to_host(alone)[0])


@pytest.mark.parametrize("n", [64, 512, 1024])
def test_invert_hadamard_ties_across_waves(gpu_device, n):
    """Sylvester Hadamard matrices: every entry is +-1, so every column's pivot
    candidates tie in magnitude across all waves (the winner selection's exact
    path), and the elimination stays in dyadic values, so the inverse H^T / n
    comes out exactly (as the oracle's blocked Gauss-Jordan gives it).  A
    winner of smaller magnitude, e.g. an eliminated zero, would show."""
    H = np.array([[1.0]])
    while H.shape[0] < n:
        H = np.block([[H, H], [H, -H]])
    d = _device.to_device(H[None], gpu_device)
    _backend.call("ticc_invert_batched", 1, n, d.data_ptr(), _stream(gpu_device))
    np.testing.assert_array_equal(_device.to_host(d)[0], H.T / n)


def test_invert_requires_row_pivoting(gpu_device):
    # zero leading diagonal: Gauss-Jordan without pivoting would divide by zero
    A = np.array([[[0.0, 1.0, 2.0], [1.0, 0.0, 3.0], [4.0, 5.0, 0.0]]])
    d = _device.to_device(A, gpu_device)
    _backend.call("ticc_invert_batched", 1, 3, d.data_ptr(), _stream(gpu_device))
    np.testing.assert_allclose(_device.to_host(d)[0], np.linalg.inv(A[0]), rtol=1e-14, atol=1e-14)


def test_riccati_bessel_matches_scipy_oracle(gpu_device):
    ell = np.array([0.0, 0.0, 1.0, 2.3, 3.0, 4.2, 10.0, 17.0, 20.3, 1.3, 0.7])
    x = np.array([1.7, 0.05, 3.0, 4.2, 516.0, 9.0, 1.5, 60.0, 300.0, 2500.0, 1.99])
    got = riccati_bessel_batch(ell, x, gpu_device.index)
    for i in range(x.size):
        ref = np.array(obessel.riccati_bessel_jy(ell[i], x[i]))
        env = max(np.hypot(ref[0], ref[1]), 1e-300)
        envp = max(np.hypot(ref[2], ref[3]), 1e-300)
        np.testing.assert_allclose(got[i, :2], ref[:2], rtol=0, atol=1e-13 * env)
        np.testing.assert_allclose(got[i, 2:], ref[2:], rtol=0, atol=1e-13 * envp)
    # order-0 elementary forms (reference test_bessel.py:7-13)
    np.testing.assert_allclose(got[0], [np.sin(1.7), -np.cos(1.7), np.cos(1.7), np.sin(1.7)], rtol=1e-14, atol=1e-14)
    # unit Wronskian at non-integer order (test_bessel.py:16-19)
    j, n, jp, npr = got[3]
    np.testing.assert_allclose(j * npr - jp * n, 1.0, rtol=1e-13, atol=1e-13)


def test_k_log_derivative_matches_scipy(gpu_device):
    nu = np.array([0.5, 1.7, 1.5, 3.5, 10.5, 48.5, 0.5, 20.5, 2.7])
    x = np.array([0.3, 3.1, 1.999, 2.0, 7.0, 2500.0, 100.0, 0.05, 476.0])
    got = k_logd_batch(nu, x, gpu_device.index)
    ref = np.array([special.kvp(a, b) / special.kv(a, b) if b < 600 else obessel.modified_bessel_K_logD(a, b) for a, b in zip(nu, x)])
    np.testing.assert_allclose(got, ref, rtol=1e-13)


def _atom_diatom_case(rng, jmax=4, Jtot=3):
    from pyticc_amd import AtomSpec, build_ScattSystem, rigid_rotor
    from pyticc_amd.basis.angle import gauss_legendre_dvr

    sysm = build_ScattSystem(AtomSpec(), rigid_rotor(1e-4, jmax), Jtot=Jtot, system_parity=1)
    ct, wt = gauss_legendre_dvr(-1.0, 1.0, 9)
    vb = prepare_atom_diatom(sysm.basis, sysm.monomer_Y.rovib, ct, wt)
    return sysm, vb


def test_contraction_matches_oracle_atom_diatom(gpu_device):
    rng = np.random.default_rng(5)
    sysm, vb = _atom_diatom_case(rng)
    N = sysm.basis.n_channel
    R = np.array([3.0, 3.5, 7.25])
    V = rng.standard_normal((R.size, 9))
    U = rng.standard_normal((N, N))
    U = U + U.T
    Ei = rng.uniform(0, 1e-3, N)
    mu = 1234.5
    W = contract_w_device(vb, R, _device.to_device(V, gpu_device), _device.to_device(U, gpu_device), _device.to_device(Ei, gpu_device), mu, gpu_device)
    blocks = [(np.asarray(vb.channel_indices[K]), vb.B_real[K], None) for K in sorted(vb.channel_indices)]
    ref = ocontract.w_base(R, mu, Ei, U, ocontract.contract(blocks, N, vb.normalization, V))
    np.testing.assert_allclose(_device.to_host(W), ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_contraction_diatom_diatom_matches_oracle_and_identity(gpu_device):
    from pyticc_amd import build_ScattSystem, rigid_rotor
    from pyticc_amd.basis.angle import gauss_legendre_dvr

    sysm = build_ScattSystem(rigid_rotor(1e-4, 2), rigid_rotor(1e-4, 3), Jtot=2, system_parity=1)
    ctX, wX = gauss_legendre_dvr(-1, 1, 6)
    ctY, wY = gauss_legendre_dvr(-1, 1, 7)
    phi, wphi = gauss_legendre_dvr(0, np.pi, 24)
    vb = prepare_diatom_diatom(sysm.basis, sysm.monomer_X.rovib, sysm.monomer_Y.rovib, ctX, wX, ctY, wY, phi, wphi)
    N = sysm.basis.n_channel
    n_ang = vb.B_real[0].shape[1]
    z = torch.zeros((N, N), dtype=torch.float64, device=gpu_device)
    z1 = torch.zeros(N, dtype=torch.float64, device=gpu_device)
    W = contract_w_device(vb, [2.0], _device.to_device(np.full((1, n_ang), 1.75), gpu_device), z, z1, 0.5, gpu_device)
    np.testing.assert_allclose(_device.to_host(W)[0], 1.75 * np.eye(N), atol=1e-12)
    rng = np.random.default_rng(8)
    V = rng.standard_normal((3, n_ang))
    W = contract_w_device(vb, [2.0, 3.0, 4.0], _device.to_device(V, gpu_device), z, z1, 0.5, gpu_device)
    blocks = [(np.asarray(vb.channel_indices[K]), vb.B_real[K], vb.B_imag[K]) for K in sorted(vb.channel_indices)]
    ref = ocontract.contract(blocks, N, vb.normalization, V)
    np.testing.assert_allclose(_device.to_host(W), ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_contraction_forms_agree(gpu_device, monkeypatch):
    """The LDS-staged contraction (16 or 4 radial points per workgroup) runs the same
    MFMA sequence per point as one wave per tile and point: W bit for bit equal; the
    forms splitting each angular sum over 4, 8 or 16 waves agree to rounding.  Radial
    points and angular points both leave ragged tails (37 points, 1008 angles)."""
    from pyticc_amd import build_ScattSystem, rigid_rotor
    from pyticc_amd.basis.angle import gauss_legendre_dvr

    sysm = build_ScattSystem(rigid_rotor(1e-4, 3), rigid_rotor(1e-4, 4), Jtot=3, system_parity=-1)
    ctX, wX = gauss_legendre_dvr(-1, 1, 6)
    ctY, wY = gauss_legendre_dvr(-1, 1, 7)
    phi, wphi = gauss_legendre_dvr(0, np.pi, 24)
    vb = prepare_diatom_diatom(sysm.basis, sysm.monomer_X.rovib, sysm.monomer_Y.rovib, ctX, wX, ctY, wY, phi, wphi)
    N = sysm.basis.n_channel
    n_ang = vb.B_real[0].shape[1]
    rng = np.random.default_rng(9)
    R = np.linspace(3.0, 9.0, 37)
    V = rng.standard_normal((R.size, n_ang))
    U = rng.standard_normal((N, N))
    U = U + U.T
    Ei = rng.uniform(0, 1e-3, N)
    dV, dU, dE = (_device.to_device(x, gpu_device) for x in (V, U, Ei))
    got = {}
    for form in ("wave", "lds1", "lds4", "split", "split8", "split16"):
        monkeypatch.setenv("TICC_CONTRACT_FORM", form)
        print(f"contraction form {form}, N={N}, n_ang={n_ang}, {R.size} points", flush=True)
        got[form] = _device.to_host(contract_w_device(vb, R, dV, dU, dE, 77.0, gpu_device))
    assert N > 100 and n_ang % 32 != 0
    np.testing.assert_array_equal(got["lds1"], got["wave"])
    np.testing.assert_array_equal(got["lds4"], got["wave"])
    blocks = [(np.asarray(vb.channel_indices[K]), vb.B_real[K], vb.B_imag[K]) for K in sorted(vb.channel_indices)]
    ref = ocontract.w_base(R, 77.0, Ei, U, ocontract.contract(blocks, N, vb.normalization, V))
    for form in ("wave", "split", "split8", "split16"):
        np.testing.assert_allclose(got[form], ref, rtol=0, atol=1e-12 * np.abs(ref).max())
