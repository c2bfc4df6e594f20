"""Per-node MLP of the learned warm start: CPU restatement of the HIP kernels
(mammos-mumag_amd/csrc/mlp.hip) for the gradient checks.  Test infrastructure only.

The network has no counterpart in the reference; this module restates this build's
own kernels so that their forward pass and gradients can be checked.  Layers:
``x (F_in) -> 128 -> 128 -> 128 -> 1``, ReLU after the three hidden layers, weights
``W_l`` stored ``[in, out]`` (FP32 masters), biases FP32.

Rounding points of the kernels (``rounded=True``): inputs, weights of layers 1-3 and
the hidden activations are rounded to bf16 (round to nearest even) before each MFMA;
products accumulate in FP32 (here: FP64); layer 4 (128 -> 1) runs in FP32 on the
bf16 activations with FP32 weights.  Backward: the activation gradients ``dz_l`` are
rounded to bf16 once; the MFMA data path, the weight-gradient GEMMs and the bias
gradients (row sums) all use the rounded values.

Pinned by: central finite differences of the unrounded FP64 network (every parameter
tensor; tests/test_oracle_mlp.py), which fixes the backward pass the kernels are
checked against.
"""
import numpy as np

WIDTH = 128


def bf16(x):
    """Round to the nearest bf16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(np.isnan(a), u, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64).reshape(np.shape(x))


def init_params(n_in, seed=0, width=WIDTH):
    """He-uniform initialisation (the kernels' host-side initialiser uses the same)."""
    rng = np.random.default_rng(seed)
    dims = [n_in, width, width, width, 1]
    p = {}
    for l in range(4):
        lim = np.sqrt(6.0 / dims[l]) if l < 3 else np.sqrt(3.0 / dims[l])
        p[f"W{l + 1}"] = rng.uniform(-lim, lim, size=(dims[l], dims[l + 1])).astype(np.float32).astype(np.float64)
        p[f"b{l + 1}"] = np.zeros(dims[l + 1])
    return p


def forward(p, X, rounded=True):
    R = bf16 if rounded else (lambda a: np.asarray(a, dtype=np.float64))
    h = R(X)
    cache = {"h0": h}
    for l in (1, 2, 3):
        z = h @ R(p[f"W{l}"]) + p[f"b{l}"]
        h = R(np.maximum(z, 0.0))
        cache[f"z{l}"] = z
        cache[f"h{l}"] = h
    y = h @ p["W4"][:, 0] + p["b4"][0]
    return y, cache


def loss_and_dy(y, t):
    n = len(y)
    d = y - t
    return float(np.mean(d * d)), 2.0 * d / n


def backward(p, cache, dy, rounded=True):
    R = bf16 if rounded else (lambda a: np.asarray(a, dtype=np.float64))
    g = {}
    h3 = cache["h3"]
    g["W4"] = (h3.T @ dy)[:, None]
    g["b4"] = np.array([dy.sum()])
    dz = np.outer(dy, p["W4"][:, 0]) * (cache["z3"] > 0)
    for l in (3, 2, 1):
        dzr = R(dz)
        g[f"b{l}"] = dzr.sum(axis=0)
        g[f"W{l}"] = cache[f"h{l - 1}"].T @ dzr
        if l > 1:
            dz = (dzr @ R(p[f"W{l}"]).T) * (cache[f"z{l - 1}"] > 0)
    return g


def features(xs, ys, zs, voxel_tag, js_of_tag, m, meas, V):
    """Node features of the structured grid (mlp.hip k_mlp_features), FP64 before the
    final bf16 rounding: m of the node and of its neighbours (-x, +x, -y, +y, -z, +z;
    0 outside the grid), coordinates / half box edge, mean magnetisation
    sum_i m_i meas_i / V, fraction of the 8 adjacent voxels that are magnetic
    (Js > 0), their mean Js / max Js, three zeros.  Node (i, j, k) = i + nx (j + ny k);
    voxel_tag is [nz-1, ny-1, nx-1]."""
    nx, ny, nz = len(xs), len(ys), len(zs)
    M = np.asarray(m, dtype=np.float64).reshape(nz, ny, nx, 3)
    P = np.zeros((nz + 2, ny + 2, nx + 2, 3))
    P[1:-1, 1:-1, 1:-1] = M
    F = np.zeros((nz, ny, nx, 32))
    c = (slice(1, -1),) * 3
    F[..., 0:3] = M
    F[..., 3:6] = P[1:-1, 1:-1, :-2]
    F[..., 6:9] = P[1:-1, 1:-1, 2:]
    F[..., 9:12] = P[1:-1, :-2, 1:-1]
    F[..., 12:15] = P[1:-1, 2:, 1:-1]
    F[..., 15:18] = P[:-2, 1:-1, 1:-1]
    F[..., 18:21] = P[2:, 1:-1, 1:-1]
    del c
    inv_half = 1.0 / max(np.abs(xs).max(), np.abs(ys).max(), np.abs(zs).max())
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    F[..., 21], F[..., 22], F[..., 23] = X * inv_half, Y * inv_half, Z * inv_half
    mb = (np.asarray(m).reshape(-1, 3) * np.asarray(meas)[:, None]).sum(axis=0) * (1.0 / V)
    F[..., 24:27] = mb
    js = np.vectorize(lambda t: js_of_tag.get(int(t), 0.0))(voxel_tag).astype(np.float64)
    Jp = np.zeros((nz + 1, ny + 1, nx + 1))
    Jp[1:-1, 1:-1, 1:-1] = js          # voxel (i-1, j-1, k-1) .. padded with air
    cnt = np.zeros((nz, ny, nx))
    tot = np.zeros((nz, ny, nx))
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                v = Jp[dk:dk + nz, dj:dj + ny, di:di + nx]
                cnt += v > 0
                tot += np.where(v > 0, v, 0.0)
    jmax = max(js_of_tag.values())
    F[..., 27] = cnt / 8.0
    F[..., 28] = np.where(cnt > 0, tot / np.maximum(cnt, 1) * (1.0 / jmax), 0.0)
    return F.reshape(-1, 32)


def to_bits(x):
    """bf16 bit patterns (as int16) of values already on the bf16 grid."""
    a = np.ascontiguousarray(bf16(x), dtype=np.float32)
    return (a.view(np.uint32) >> 16).astype(np.uint16).view(np.int16)


def from_bits(b):
    return (np.asarray(b).view(np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
