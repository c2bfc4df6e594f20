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


def test_features_brute_force():
    """oracle mlp.features (vectorised) against a per-node loop over the definition on a
    small non-cubic grid with random m and voxel materials: the neighbour order (-x, +x,
    -y, +y, -z, +z), zeros outside the grid, coordinates over the half box edge, the mean
    magnetisation, and the 8 adjacent voxels (indices i-1..i, j-1..j, k-1..k) with their
    magnetic fraction and mean Js over the largest Js."""
    from oracle import mlp as omlp
    rng = np.random.default_rng(11)
    nx, ny, nz = 5, 4, 6
    xs = np.cumsum(rng.uniform(0.5, 2.0, nx)); xs -= xs.mean()
    ys = np.cumsum(rng.uniform(0.5, 2.0, ny)); ys -= ys.mean()
    zs = np.cumsum(rng.uniform(0.5, 2.0, nz)); zs -= zs.mean()
    tag = rng.integers(0, 4, size=(nz - 1, ny - 1, nx - 1))
    js_of_tag = {1: 1.61, 2: 0.8, 3: 1.2}          # tag 0: air (Js 0)
    n = nx * ny * nz
    m = rng.normal(size=(n, 3))
    meas = rng.uniform(0.0, 2.0, n)
    V = 17.0
    F = omlp.features(xs, ys, zs, tag, js_of_tag, m.ravel(), meas, V)
    half = max(np.abs(xs).max(), np.abs(ys).max(), np.abs(zs).max())
    mb = (m * meas[:, None]).sum(axis=0) / V
    jmax = max(js_of_tag.values())

    def node(i, j, k):
        return i + nx * (j + ny * k)

    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                f = F[node(i, j, k)]
                assert np.array_equal(f[0:3], m[node(i, j, k)])
                for q, (di, dj, dk) in enumerate([(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]):
                    a, b, c = i + di, j + dj, k + dk
                    want = m[node(a, b, c)] if (0 <= a < nx and 0 <= b < ny and 0 <= c < nz) else np.zeros(3)
                    assert np.array_equal(f[3 + 3 * q:6 + 3 * q], want), (i, j, k, q)
                assert np.allclose(f[21:24], [xs[i] / half, ys[j] / half, zs[k] / half], rtol=0, atol=1e-15)
                assert np.allclose(f[24:27], mb, rtol=1e-14, atol=0)
                vals = [js_of_tag.get(int(tag[c, b, a]), 0.0)
                        for c in (k - 1, k) for b in (j - 1, j) for a in (i - 1, i)
                        if 0 <= a < nx - 1 and 0 <= b < ny - 1 and 0 <= c < nz - 1]
                mag = [v for v in vals if v > 0]
                assert f[27] == len(mag) / 8.0
                assert np.isclose(f[28], (sum(mag) / len(mag) / jmax) if mag else 0.0, rtol=1e-14, atol=0)
                assert np.all(f[29:32] == 0)
