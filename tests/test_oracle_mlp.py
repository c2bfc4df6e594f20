"""The MLP oracle's backward pass against central finite differences (unrounded FP64),
and its bf16 rounding against the definition (round to nearest, ties to even)."""
import numpy as np

from oracle import mlp


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 0.0, 1e-30, 3.0e38])
    r = mlp.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 3 2^-8 rounds up to 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2.0 ** -6 and r[3] == 1.0 + 2.0 ** -7
    assert r[4] == -2.5 and r[5] == 0.0
    # 7 stored mantissa bits: relative error at most half an ulp, 2^-8
    y = np.random.default_rng(0).normal(size=1000) * 10.0 ** np.random.default_rng(1).uniform(-5, 5, 1000)
    assert np.all(np.abs(mlp.bf16(y) - y) <= 2.0 ** -8 * np.abs(y) * (1 + 1e-6))


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    n_in, n = 32, 40
    p = mlp.init_params(n_in, seed=5, width=16)
    for k in p:
        if k.startswith("b"):
            p[k] = rng.normal(scale=0.1, size=p[k].shape)
    X = rng.normal(size=(n, n_in))
    t = rng.normal(size=n)

    def L(q):
        y, _ = mlp.forward(q, X, rounded=False)
        return mlp.loss_and_dy(y, t)[0]

    y, cache = mlp.forward(p, X, rounded=False)
    _, dy = mlp.loss_and_dy(y, t)
    g = mlp.backward(p, cache, dy, rounded=False)
    eps = 1e-6
    for k in p:
        idx = [tuple(rng.integers(0, s) for s in p[k].shape) for _ in range(6)]
        for ix in idx:
            q = {kk: vv.copy() for kk, vv in p.items()}
            q[k][ix] += eps
            lp = L(q)
            q[k][ix] -= 2 * eps
            lm = L(q)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[k][ix]) <= 1e-6 * max(1.0, abs(fd)), (k, ix, fd, g[k][ix])


def test_rounded_forward_close_to_exact():
    rng = np.random.default_rng(4)
    p = mlp.init_params(32, seed=1)
    X = rng.normal(size=(64, 32))
    y0, _ = mlp.forward(p, X, rounded=False)
    y1, _ = mlp.forward(p, X, rounded=True)
    assert np.linalg.norm(y1 - y0) / np.linalg.norm(y0) < 2e-2
