"""The learned warm start's network on the GPU against the MLP oracle (oracle/mlp.py):
bf16 MFMA forward and the gradients of the hand-written backward kernels, the feature
kernel, training, and the ablation — a loop with and without the learned warm start
reaches the same answer, and the test reports the PCG iterations each needs.

Tolerances: bf16 activations make single roundings flip with the FP32 summation order
of the MFMA, so the forward is compared to 1e-2 and the gradients to 3e-2 (relative
2-norm per tensor) against the oracle evaluated with the same rounding points.
"""
import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def _flat_to_dict(p):
    d, o = {}, 0
    for name, shape in (("W1", (32, 128)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)),
                        ("W3", (128, 128)), ("b3", (128,)), ("W4", (128, 1)), ("b4", (1,))):
        n = int(np.prod(shape))
        d[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    assert o == len(p)
    return d


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def mesh():
    from mammos_mumag_amd import generate
    return generate.build(edge=24.0, n=8, layers=3, grains=4, cone=15.0, seed=7)


@pytest.fixture(scope="module")
def model(mesh):
    from mammos_mumag_amd import Model
    print("[gpu] mlp: structured context", flush=True)
    return Model.from_structured(mesh)


def _batch(n, rng):
    import torch
    X = rng.normal(size=(n, 32)) * 0.7
    X[:, 29:] = 0.0
    Xb = omlp.to_bits(X)
    return omlp.from_bits(Xb), torch.tensor(Xb, device="cuda")


def test_forward_matches_oracle(model):
    from mammos_mumag_amd.learned import Network
    print("[gpu] mlp forward", flush=True)
    rng = np.random.default_rng(1)
    net = Network(model, max_batch=4096, seed=11)
    p = rng.normal(size=37377).astype(np.float32) * 0.1
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    X, Xd = _batch(1000, rng)
    y = net.forward(Xd).cpu().numpy()
    y_ref, _ = omlp.forward(_flat_to_dict(p), X, rounded=True)
    assert _rel(y, y_ref) < 1e-2
    net.close()


@pytest.mark.parametrize("n", [200, 1024])
def test_gradients_match_oracle(model, n):
    import torch
    from mammos_mumag_amd.learned import Network
    print(f"[gpu] mlp gradients, batch {n}", flush=True)
    rng = np.random.default_rng(2 + n)
    net = Network(model, max_batch=2048, seed=5)
    p = net.get_params()
    P = _flat_to_dict(p)
    for k in ("b1", "b2", "b3"):
        P[k] = rng.normal(scale=0.05, size=P[k].shape)
    flat = np.concatenate([P[k].ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4")])
    net.set_params(flat.astype(np.float32))
    P = _flat_to_dict(flat.astype(np.float32))
    X, Xd = _batch(n, rng)
    t = rng.normal(size=n)
    g, loss = net.grads(Xd, torch.tensor(t, dtype=torch.float32, device="cuda"))
    y_ref, cache = omlp.forward(P, X, rounded=True)
    loss_ref, dy = omlp.loss_and_dy(y_ref, t)
    g_ref = omlp.backward(P, cache, dy, rounded=True)
    G = _flat_to_dict(g)
    assert abs(loss / loss_ref - 1) < 1e-2
    for k in g_ref:
        assert _rel(G[k], g_ref[k]) < 3e-2, (k, _rel(G[k], g_ref[k]))
    net.close()


def test_features_match_oracle(model, mesh):
    import torch
    from mammos_mumag_amd.learned import features
    print("[gpu] mlp features", flush=True)
    rng = np.random.default_rng(4)
    m = rng.normal(size=(model.N, 3))
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    m *= (model.meas > 0)[:, None]
    X = features(model, model.to_device(m.ravel())).cpu().numpy()
    F = omlp.features(mesh.xs, mesh.ys, mesh.zs, mesh.voxel_tag,
                      {t: p["Js"] for t, p in mesh.tags_props.items()}, m.ravel(), model.meas, model.V)
    got, ref = omlp.from_bits(X), omlp.bf16(F)
    local = [c for c in range(32) if c not in (24, 25, 26)]
    assert np.array_equal(got[:, local], ref[:, local])
    # the mean magnetisation is a sum in another order: equal up to one bf16 rounding
    assert np.all(np.abs(got[:, 24:27] - ref[:, 24:27]) <= 2.0 ** -7 * np.abs(ref[:, 24:27]) + 1e-30)
    del torch


def test_training_reduces_loss(model):
    import torch
    from mammos_mumag_amd.learned import Network, train
    print("[gpu] mlp training on a smooth target", flush=True)
    rng = np.random.default_rng(6)
    X, Xd = _batch(8192, rng)
    t = np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 21]
    net = Network(model, max_batch=4096, seed=3)
    hist = train(net, Xd, torch.tensor(t, dtype=torch.float32, device="cuda"), steps=400, batch=4096,
                 lr=2e-3, log_every=100)
    print("[gpu] loss history", hist, flush=True)
    assert hist[-1][1] < 0.1 * hist[0][1]
    net.close()


def test_warm_start_ablation(mesh):
    """Train on (m, u) pairs of another microstructure's loop, then run the same field
    steps on this one without the learned warm start, with it before every solve, and
    with the measured policy (per solve the faster form): the answers agree to the solver
    tolerance; the iteration counts and solve times show what the network gains."""
    import torch
    from mammos_mumag_amd import Model, _lib, generate
    from mammos_mumag_amd.learned import Network, collect, train
    print("[gpu] warm-start ablation: training data", flush=True)
    p = generate.loop_params("cube40_single")
    other = generate.build(edge=24.0, n=8, layers=3, grains=4, cone=15.0, seed=99)
    src = Model.from_structured(other)
    X, u = collect(src, p, steps=40, nodes_per_step=2048, hstep=-0.4)
    net = Network(src, max_batch=8192, seed=1)
    hist = train(net, X, u, steps=600, batch=8192, lr=2e-3)
    params, scale = net.get_params(), net.scale
    print("[gpu] trained: loss", hist[0][1], "->", hist[-1][1], flush=True)
    src.close()

    results = {}
    for learned in (False, True, "measured"):
        model = Model.from_structured(mesh)
        if learned:
            n2 = Network(model, max_batch=32, seed=0)
            n2.set_params(params)
            n2.scale = scale
            model.warm_start(n2)
            if learned == "measured":
                model.warm_start_policy("measured")
        m = model.to_device(model.uniform_m(p["m"]))
        m_ref, uu, gF = m.clone(), model.empty1(), model.empty3()
        st = _lib.Stats()
        hext, rec = p["hstart"], []
        for k in range(12):
            E = model.minimize(m, hext, p["h"], p, m_ref, uu, gF, st)
            rec.append((E, model.compute_mh(m, p["h"])))
            hext += -0.4
        results[learned] = (np.array(rec), st.as_dict(),
                            dict(model.warm_start_stats(), **model.warm_start_report()) if learned else None)
        model.close()
    (r0, s0, _), (r1, s1, ws) = results[False], results[True]
    (r2, s2, wm) = results["measured"]
    print(f"[gpu] measured policy: {wm['calls']} of {wm['solves']} solves warm-started, mean "
          f"{1e3 * wm['mean_s_plain']:.2f} ms plain / {1e3 * wm['mean_s_warm']:.2f} ms warm-started", flush=True)
    # every solve is timed; both forms are measured at least 3 times
    assert wm["solves"] == s2["pcg_calls"] and wm["calls"] == wm["warm_solves"]
    assert 3 <= wm["warm_solves"] <= wm["solves"] - 3
    assert np.max(np.abs(r2[:, 1] - r0[:, 1])) < 1e-6
    assert np.max(np.abs(r2[:, 0] / r0[:, 0] - 1)) < 1e-8
    print(f"[gpu] plain PCG iterations {s0['pcg_iters']}, learned warm start {s1['pcg_iters']} "
          f"({ws['calls']} warm-started solves); NCG {s0['ncg_iters']} / {s1['ncg_iters']}", flush=True)
    assert ws["calls"] == s1["pcg_calls"]
    assert np.max(np.abs(r1[:, 1] - r0[:, 1])) < 1e-6
    assert np.max(np.abs(r1[:, 0] / r0[:, 0] - 1)) < 1e-8
