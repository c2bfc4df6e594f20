"""The learned preconditioner (mm_lp_*) on a small generated structured mesh.

* Residual features (mm_lp_features) against their definition restated with numpy (the
  Jacobi-scaled residual of the node and of its neighbours one and two nodes away along
  each axis over its RMS, the spacings over their mean, the face flag, the magnetic
  fraction, the coordinates): equal to one bf16 rounding.
* The CG driver of mm_lp_pcg without the network is the Jacobi PCG: the same iteration
  count as mm_pcg (within one) and the same solution (1e-9).
* With a network trained here on residual samples of another microstructure, every
  variant ends at the Jacobi PCG's tolerance with its solution (the network's Galerkin
  steps change the iteration count, never the answer); the counts are printed — the gain
  or loss against the Jacobi PCG is what the network buys.
"""
import numpy as np
import pytest

from mammos_mumag_amd import generate

pytestmark = pytest.mark.gpu


def _bf16(a):
    """Round float32 to bfloat16 (ties to even) and back."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _features(mesh, model, r, invd):
    nx, ny, nz = mesh.shape
    N = nx * ny * nz
    w = (r * invd).reshape(nz, ny, nx)
    s = np.sqrt(np.mean(w ** 2))
    f = np.zeros((nz, ny, nx, 32))

    def shifted(di, dj, dk):
        out = np.zeros_like(w)
        src = w[max(dk, 0):nz + min(dk, 0), max(dj, 0):ny + min(dj, 0), max(di, 0):nx + min(di, 0)]
        out[max(-dk, 0):nz - max(dk, 0), max(-dj, 0):ny - max(dj, 0), max(-di, 0):nx - max(di, 0)] = src
        return out / s
    offs = [(0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1),
            (-2, 0, 0), (2, 0, 0), (0, -2, 0), (0, 2, 0), (0, 0, -2), (0, 0, 2)]
    for q, o in enumerate(offs):
        f[..., q] = shifted(*o)
    def pads(x):
        h = np.diff(x)
        return np.concatenate([[0.0], h]), np.concatenate([h, [0.0]])   # left, right spacing per node
    hxl, hxu = pads(mesh.xs)
    hyl, hyu = pads(mesh.ys)
    hzl, hzu = pads(mesh.zs)
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    H = np.stack([hxl[X], hxu[X], hyl[Y], hyu[Y], hzl[Z], hzu[Z]], axis=-1)
    cnt = (H > 0).sum(axis=-1, keepdims=True)
    f[..., 13:19] = H * cnt / H.sum(axis=-1, keepdims=True)
    face = (X == 0) | (Y == 0) | (Z == 0) | (X == nx - 1) | (Y == ny - 1) | (Z == nz - 1)
    f[..., 19] = face
    vt = mesh.voxel_tag
    magv = np.vectorize(lambda t: mesh.tags_props[int(t)]["Js"] > 0)(vt).astype(float)
    pad = np.zeros((nz + 1, ny + 1, nx + 1))
    pad[1:nz, 1:ny, 1:nx] = magv
    frac = sum(pad[dk:dk + nz, dj:dj + ny, di:di + nx] for dk in (0, 1) for dj in (0, 1) for di in (0, 1)) / 8
    f[..., 20] = frac
    half = max(np.abs(mesh.xs).max(), np.abs(mesh.ys).max(), np.abs(mesh.zs).max())
    f[..., 21] = mesh.xs[X] / half
    f[..., 22] = mesh.ys[Y] / half
    f[..., 23] = mesh.zs[Z] / half
    return f.reshape(N, 32).astype(np.float32)


@pytest.fixture(scope="module")
def mesh():
    return generate.build(edge=24.0, n=8, layers=3, grains=4, cone=15.0, seed=7)


def test_residual_features(mesh):
    import ctypes
    import torch
    from mammos_mumag_amd import Model, _lib
    from mammos_mumag_amd.learned import lp_features
    model = Model.from_structured(mesh)
    model.lp_setup()
    dex, dk = np.empty(3 * model.N), np.empty(model.N)
    _lib.check(model._lib.mm_get_diag_host(model._ctx, dex.ctypes.data_as(_lib.c_double_p),
                                           dk.ctypes.data_as(_lib.c_double_p)), "diag")
    rng = np.random.default_rng(4)
    r = rng.normal(size=model.N)
    X = lp_features(model, model.to_device(r)).cpu().numpy().view(np.uint16)
    got = (X.astype(np.uint32) << 16).view(np.float32)
    ref = _features(mesh, model, r, 1.0 / dk)
    refb = _bf16(ref)
    ulp = np.abs(refb) * 2.0 ** -7
    assert np.all(np.abs(got - refb) <= ulp + 1e-30)
    assert np.mean(got == refb) > 0.999
    model.close()
    del ctypes, torch


def test_learned_cg_keeps_the_answer(mesh):
    import torch
    from mammos_mumag_amd import Model
    from mammos_mumag_amd.learned import Network, collect_both, train
    p = generate.loop_params("cube40_single")
    print("[gpu] learned preconditioner: residual samples of another microstructure", flush=True)
    other = Model.from_structured(generate.build(edge=24.0, n=8, layers=3, grains=4, cone=15.0, seed=99))
    _, _, Xl, tl = collect_both(other, p, steps=6, hstep=-1.0, nodes_ws=64, nodes_lp=2048)
    net = Network(other, max_batch=8192, seed=1)
    hist = train(net, Xl, tl, steps=400, batch=8192, lr=2e-3)
    params, scale = net.get_params(), net.scale
    print(f"[gpu] trained: loss {hist[0][1]:.3e} -> {hist[-1][1]:.3e}", flush=True)
    other.close()

    model = Model.from_structured(mesh)
    model.lp_setup()
    n2 = Network(model, max_batch=32, seed=0)
    n2.set_params(params)
    n2.scale = scale
    m = model.to_device(model.uniform_m([0.0, 0.0, 1.0]))
    b = model.divergence(m)
    x_ref = model.empty1()
    it_ref, rc = model.pcg(b, x_ref, 1e-10)
    assert rc == 0
    x = model.empty1()
    it0, apps0, rc0 = model.lp_pcg(b, x, 1e-10, net=None)
    assert rc0 == 0 and apps0 == 0 and abs(it0 - it_ref) <= 1
    assert float(torch.max(torch.abs(x - x_ref)) / torch.max(torch.abs(x_ref))) < 1e-9
    counts = {}
    for every in (1000000, 64, 16):
        x = model.empty1()
        it, apps, rc = model.lp_pcg(b, x, 1e-10, maxiter=20000, net=n2, every=every)
        counts[every] = (it, apps)
        assert rc == 0 and apps >= 1
        assert float(torch.max(torch.abs(x - x_ref)) / torch.max(torch.abs(x_ref))) < 1e-8
    print(f"[gpu] Jacobi PCG {it_ref} iterations; with the network's Galerkin steps "
          + ", ".join(f"every {e}: {c[0]} iterations ({c[1]} steps)" for e, c in counts.items()), flush=True)
    model.close()
