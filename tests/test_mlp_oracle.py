"""The MLP oracle's backward pass against central finite differences (FP64, no rounding), and its
BF16 operand rounding against known values.  No reference counterpart exists."""
import numpy as np
import pytest

import mlp_oracle as M


def small_net(seed=0, dims=(5, 7, 7, 3)):
    rng = np.random.default_rng(seed)
    return [(rng.standard_normal((do, di)) / np.sqrt(di), rng.standard_normal(do) * 0.1)
            for di, do in zip(dims[:-1], dims[1:])]


@pytest.mark.parametrize("weighted", [False, True])
def test_gradients_match_finite_differences(weighted):
    layers = small_net()
    rng = np.random.default_rng(1)
    x = rng.standard_normal((9, 5))
    t = rng.standard_normal((9, 3))
    w = rng.uniform(0.0, 2.0, (9, 3)) if weighted else None
    _, g = M.loss_grads(layers, x, t, bf16_ops=False, w=w)
    h = 1e-6
    for l, (W, b) in enumerate(layers):
        for arr, ga in ((W, g[l][0]), (b, g[l][1])):
            flat = arr.reshape(-1)
            gflat = ga.reshape(-1)
            for i in range(0, flat.size, 3):
                old = flat[i]
                flat[i] = old + h
                lp = M.loss_grads(layers, x, t, bf16_ops=False, w=w)[0]
                flat[i] = old - h
                lm = M.loss_grads(layers, x, t, bf16_ops=False, w=w)[0]
                flat[i] = old
                fd = (lp - lm) / (2 * h)
                assert abs(fd - gflat[i]) <= 1e-7 * max(1.0, abs(fd)), (l, i, fd, gflat[i])


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, 3.0e38, -2.5, 1.0 + 3 * 2 ** -9])
    # 1+2^-8 is a tie between 1 and 1+2^-7: rounds to even (1.0); 1+2^-7+2^-9 rounds down
    expect = np.array([1.0, 1.0, 1.0 + 2 ** -7, 2.9975e38, -2.5, 1.0 + 2 ** -7])
    got = M.bf16(vals)
    assert np.allclose(got[[0, 1, 2, 4, 5]], expect[[0, 1, 2, 4, 5]], rtol=0, atol=0)
    assert abs(got[3] / 3.0e38 - 1) < 2 ** -8
