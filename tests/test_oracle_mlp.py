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
