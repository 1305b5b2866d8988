"""The surrogate's oracle network: its backward pass against finite differences of its
forward pass (float64, no rounding), and its BF16 rounding against the definition."""

import numpy as np

from oracle import mlp


def _small(rng, dims=(5, 7, 6, 1)):
    Ws = [rng.standard_normal((dout, din)) * 0.5 for din, dout in zip(dims[:-1], dims[1:])]
    bs = [rng.standard_normal(dout) * 0.1 for dout in dims[1:]]
    return Ws, bs


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    Ws, bs = _small(rng)
    F = rng.standard_normal((9, 5))
    T = rng.standard_normal(9)
    loss, gW, gb = mlp.loss_and_grads(F, T, Ws, bs, round_bf16=False)
    h = 1e-6
    for l in range(len(Ws)):
        r, c = Ws[l].shape
        for idx in [(0, 0), (min(1, r - 1), min(2, c - 1)), (r - 1, c - 1)]:
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            num = (mlp.loss_and_grads(F, T, Wp, bs, round_bf16=False)[0] - mlp.loss_and_grads(F, T, Wm, bs, round_bf16=False)[0]) / (2 * h)
            np.testing.assert_allclose(gW[l][idx], num, rtol=1e-6, atol=1e-9)
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][0] += h
        bm[l][0] -= h
        num = (mlp.loss_and_grads(F, T, Ws, bp, round_bf16=False)[0] - mlp.loss_and_grads(F, T, Ws, bm, round_bf16=False)[0]) / (2 * h)
        np.testing.assert_allclose(gb[l][0], num, rtol=1e-6, atol=1e-9)


def test_bf16_rounding_definition():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7 + 2**-9, -3.14159, 1e-30, 65504.0], dtype=np.float32)
    got = mlp.bf16(x)
    np.testing.assert_array_equal(got[:2], [1.0, 1.0])             # tie to even
    assert got[2] == 1.0 + 2**-6 or got[2] == 1.0 + 2**-7           # rounds to a representable neighbour
    assert abs(got[3] + 3.140625) < 1e-12                           # 8 significant bits
    np.testing.assert_allclose(got, x, rtol=2**-8)


def test_silu_derivative():
    z = np.linspace(-6, 6, 41)
    h = 1e-6
    np.testing.assert_allclose(mlp.dsilu(z), (mlp.silu(z + h) - mlp.silu(z - h)) / (2 * h), rtol=1e-7, atol=1e-9)


def test_basis_head_backward_matches_finite_differences():
    """The angular-basis head (prediction = output . G of the sample) against
    finite differences, and it reduces to the scalar head for G = e_0."""
    rng = np.random.default_rng(3)
    Ws, bs = _small(rng, dims=(6, 7, 5, 4))
    F = rng.standard_normal((11, 6))
    G = rng.standard_normal((11, 4))
    T = rng.standard_normal(11)
    loss, gW, gb = mlp.basis_loss_and_grads(F, G, T, Ws, bs, round_bf16=False)
    h = 1e-6
    for l in range(len(Ws)):
        r, c = Ws[l].shape
        for idx in [(0, 0), (r - 1, c - 1)]:
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            num = (mlp.basis_loss_and_grads(F, G, T, Wp, bs, round_bf16=False)[0]
                   - mlp.basis_loss_and_grads(F, G, T, Wm, bs, round_bf16=False)[0]) / (2 * h)
            np.testing.assert_allclose(gW[l][idx], num, rtol=1e-6, atol=1e-9)
    e0 = np.zeros_like(G)
    e0[:, 0] = 1.0
    a = mlp.basis_loss_and_grads(F, e0, T, Ws, bs, round_bf16=False)
    b = mlp.loss_and_grads(F, T, Ws, bs, round_bf16=False)
    assert a[0] == b[0]
    for x, y in zip(a[1] + a[2], b[1] + b[2]):
        np.testing.assert_allclose(x, y, rtol=1e-14, atol=1e-15)


def test_radial_features():
    R = np.array([4.0, 8.0, 4.7, 23.3])
    f = mlp.radial_features(R, 4.0)
    np.testing.assert_allclose(f[:2, :3], [[1.0, 1.0, 1.0], [0.5, 0.25, 0.125]])
    np.testing.assert_array_equal(f[:2, 3:], 0.0)  # powers of two round exactly
    # the BF16 inputs and their residuals resolve x = 4/R to about 16 bits
    x = 4.0 / R
    approx = mlp.bf16(f[:, 0].astype(np.float32)) + mlp.bf16(f[:, 3].astype(np.float32)) / 256.0
    assert np.max(np.abs(approx - x) / x) < 2.0**-15
    assert np.max(np.abs(mlp.bf16(f[:, 0].astype(np.float32)) - x) / x) > 2.0**-12


def test_split_bf16_forward_is_near_the_unrounded_network():
    """hi*hi + hi*lo + lo*hi of BF16 parts keeps about 16 bits per product: the split
    forward sits orders of magnitude closer to the float64 network than the plain
    BF16 forward (the features enter rounded to BF16 in both, so they are
    pre-rounded here), and the low part of x is what BF16 rounding drops."""
    rng = np.random.default_rng(11)
    Ws, bs = _small(rng, dims=(6, 64, 64, 64, 8))
    F = mlp.bf16(rng.standard_normal((200, 6)).astype(np.float32))
    exact = mlp.forward(F, Ws, bs, round_bf16=False)[1][1][-1]
    plain = mlp.forward(F, Ws, bs)[1][1][-1]
    split = mlp.forward(F, Ws, bs, split=True)[1][1][-1]
    scale = np.abs(exact).max()
    e_plain = np.abs(plain - exact).max() / scale
    e_split = np.abs(split - exact).max() / scale
    assert e_plain > 1e-3 and e_split < 2e-5 and e_split < e_plain / 100, (e_plain, e_split)
    x = rng.standard_normal(1000).astype(np.float32)
    hi, lo = mlp.bf16(x), mlp.bf16_lo(x)
    assert np.all(np.abs(x - hi - lo) <= np.abs(x) * 2.0**-16)
