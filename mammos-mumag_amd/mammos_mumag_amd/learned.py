"""Learned warm start of the potential solve (no counterpart in the reference).

A per-node MLP (the ``mm_mlp_*`` entry points of ``libmumag_hip.so``: bf16 MFMA forward
and backward kernels, FP32 master weights, Adam) maps the features of a node and its
1-ring (magnetisation of the node and its six axis neighbours, coordinates, the mean
magnetisation, the magnetic fraction and mean Js of the adjacent voxels) to the scalar
potential u.  It is trained on the GPU on (m, u) pairs that the plain FP64 solver
generates along hysteresis loops of other seeded microstructures, and used through
``Model.warm_start``: before every potential solve, u moves by the Galerkin step along
the change of the prediction, then the Jacobi PCG finishes to the unchanged tolerance —
the network changes the iteration count, never the answer.

Structured (Kuhn grid) contexts only.
"""
import ctypes
import math
import time

import numpy as np

from . import _lib
from .model import _ptr, _torch

N_IN = _lib.MLP_IN
N_PARAMS = _lib.MLP_PARAMS


class Network:
    """The per-node MLP on ``model``'s device; training batches up to ``max_batch`` nodes."""

    def __init__(self, model, max_batch=65536, seed=0):
        self.model = model
        self._lib = model._lib
        self.max_batch = int(max_batch)
        self._net = ctypes.c_void_p()
        _lib.check(self._lib.mm_mlp_create(model._ctx, self.max_batch, int(seed), ctypes.byref(self._net)),
                   "mm_mlp_create")
        model._attach(self)
        self.scale = 1.0    # u = scale * output
        self.steps = 0

    def close(self):
        if getattr(self, "_net", None) and self._net.value:
            self._lib.mm_mlp_destroy(self._net)
            self._net = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------------------- parameters
    def get_params(self):
        p = np.empty(N_PARAMS, dtype=np.float32)
        _lib.check(self._lib.mm_mlp_get_params(self._net, p.ctypes.data_as(_lib.c_float_p)), "mm_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != N_PARAMS:
            raise ValueError(f"expected {N_PARAMS} parameters, got {p.size}")
        _lib.check(self._lib.mm_mlp_set_params(self._net, p.ctypes.data_as(_lib.c_float_p)), "mm_mlp_set_params")

    def save(self, path, **meta):
        np.savez(path, params=self.get_params(), scale=np.float64(self.scale), **meta)

    def load(self, path):
        z = np.load(path, allow_pickle=False)
        self.set_params(z["params"])
        self.scale = float(z["scale"])
        return {k: z[k] for k in z.files if k not in ("params", "scale")}

    # -------------------------------------------------------------- kernels
    def forward(self, X, y=None):
        torch = _torch()
        n = X.shape[0]
        y = torch.empty(n, dtype=torch.float32, device=X.device) if y is None else y
        self.model._sync()
        _lib.check(self._lib.mm_mlp_forward(self._net, n, _ptr(X, n * N_IN, "int16"), _ptr(y, n, "float32")),
                   "mm_mlp_forward")
        return y

    def grads(self, X, t):
        n = X.shape[0]
        g = np.empty(N_PARAMS, dtype=np.float32)
        loss = ctypes.c_double()
        self.model._sync()
        _lib.check(self._lib.mm_mlp_grads(self._net, n, _ptr(X, n * N_IN, "int16"), _ptr(t, n, "float32"),
                                          g.ctypes.data_as(_lib.c_float_p), ctypes.byref(loss)), "mm_mlp_grads")
        return g, loss.value

    def train_step(self, X, t, lr, want_loss=False):
        n = X.shape[0]
        loss = ctypes.c_double()
        self.model._sync()   # the batch is made on torch's stream
        _lib.check(self._lib.mm_mlp_train_step(self._net, n, _ptr(X, n * N_IN, "int16"), _ptr(t, n, "float32"),
                                               float(lr), ctypes.byref(loss) if want_loss else None),
                   "mm_mlp_train_step")
        self.steps += 1
        return loss.value if want_loss else None


def features(model, m, X=None):
    """bf16 feature rows (as int16) of every node for magnetisation m (device f64[3N])."""
    torch = _torch()
    X = torch.empty((model.N, N_IN), dtype=torch.int16, device=m.device) if X is None else X
    model._sync()
    _lib.check(model._lib.mm_mlp_features(model._ctx, _ptr(m, 3 * model.N), _ptr(X, model.N * N_IN, "int16")),
               "mm_mlp_features")
    return X


def collect(model, params, steps, nodes_per_step, seed=0, hstep=None, log=None):
    """(X, u) samples of the plain solver along a loop of ``steps`` field steps on
    ``model`` (the field sweep of ``params``, optionally with a coarser ``hstep``):
    ``nodes_per_step`` random nodes of every converged field step."""
    torch = _torch()
    from . import _lib as L
    gen = torch.Generator(device=f"cuda:{model.device}")
    gen.manual_seed(int(seed))
    m = model.to_device(model.uniform_m(params["m"]))
    m_ref = m.clone()
    u = model.empty1()
    gF = model.empty3()
    X = torch.empty((model.N, N_IN), dtype=torch.int16, device=m.device)
    Xs, us = [], []
    hext = params["hstart"]
    dh = params["hstep"] if hstep is None else hstep
    st = L.Stats()
    for k in range(steps):
        model.minimize(m, hext, params["h"], params, m_ref, u, gF, st)
        features(model, m, X)
        idx = torch.randint(0, model.N, (nodes_per_step,), generator=gen, device=m.device)
        Xs.append(X[idx].clone())
        us.append(u[idx].float())
        if log:
            log(f"[collect] step {k}: h = {hext:.2f} T, M.h = {model.compute_mh(m, params['h']):.4f}")
        hext += dh
    return torch.cat(Xs), torch.cat(us)


def train(net, X, u, steps=2000, batch=16384, lr=1e-3, lr_min=1e-5, seed=0, log=None, log_every=200):
    """Adam on mean((net(X) - u / scale)^2) with a cosine learning-rate schedule;
    ``net.scale`` is set to the standard deviation of u.  Returns the loss history."""
    torch = _torch()
    net.scale = float(u.double().std().item()) or 1.0
    t = (u / net.scale).float().contiguous()
    gen = torch.Generator(device=X.device)
    gen.manual_seed(int(seed))
    batch = min(batch, net.max_batch, X.shape[0])
    hist = []
    t0 = time.perf_counter()
    for s in range(steps):
        idx = torch.randint(0, X.shape[0], (batch,), generator=gen, device=X.device)
        lr_s = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * s / max(steps - 1, 1)))
        want = (s % log_every == 0) or s == steps - 1
        loss = net.train_step(X[idx].contiguous(), t[idx].contiguous(), lr_s, want_loss=want)
        if want:
            hist.append((s, loss))
            if log:
                log(f"[train] step {s}: loss {loss:.4e} (lr {lr_s:.2e}, {time.perf_counter() - t0:.1f} s)")
    return hist


def train_for_config(config, seeds, steps, hstep, nodes_per_step, train_steps, batch, lr, out,
                     log=lambda msg: print(msg, flush=True)):
    """Train the warm-start network for a generated workload: (m, u) pairs from
    plain-FP64 loops of the microstructures ``seeds`` (none of them the evaluation
    seed), then Adam on the GPU; saves parameters, output scale and provenance."""
    torch = _torch()
    from .generate import build_config, loop_params
    from .model import Model
    p = loop_params(config)
    Xs, us = [], []
    t0 = time.perf_counter()
    for sd in seeds:
        g = build_config(config, seed=sd)
        model = Model.from_structured(g)
        X, u = collect(model, p, steps, nodes_per_step, seed=sd, hstep=hstep,
                       log=lambda msg: log(f"{msg} (seed {sd}, {time.perf_counter() - t0:.1f} s)"))
        Xs.append(X)
        us.append(u)
        model.close()
        log(f"[learned] seed {sd}: {X.shape[0]} samples ({time.perf_counter() - t0:.1f} s)")
    X = torch.cat(Xs)
    u = torch.cat(us)
    g = build_config(config, seed=seeds[0])
    model = Model.from_structured(g)
    net = Network(model, max_batch=batch, seed=0)
    hist = train(net, X, u, steps=train_steps, batch=batch, lr=lr, log=log, log_every=max(train_steps // 10, 1))
    net.save(out, config=config, seeds=np.asarray(seeds), steps=steps, hstep=hstep,
             nodes_per_step=nodes_per_step, train_steps=train_steps, batch=batch, lr=lr,
             loss_first=hist[0][1], loss_last=hist[-1][1])
    log(f"[learned] saved {out}: loss {hist[0][1]:.3e} -> {hist[-1][1]:.3e} ({time.perf_counter() - t0:.1f} s)")
    model.close()
    return hist


if __name__ == "__main__":
    import argparse
    import pathlib
    ap = argparse.ArgumentParser(description="train the learned warm start for a generated workload")
    ap.add_argument("--config", default="voronoi64_250nm")
    ap.add_argument("--seeds", type=int, nargs="+", default=[2001, 2002, 2003])
    ap.add_argument("--steps", type=int, default=24, help="field steps per loop")
    ap.add_argument("--hstep", type=float, default=-0.667, help="field step (T) of the training loops")
    ap.add_argument("--nodes", type=int, default=131072, help="sampled nodes per field step")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=32768)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    out = a.out or str(pathlib.Path(__file__).resolve().parent / "data" / f"warmstart_{a.config}.npz")
    pathlib.Path(out).parent.mkdir(exist_ok=True)
    train_for_config(a.config, a.seeds, a.steps, a.hstep, a.nodes, a.train_steps, a.batch, a.lr, out)
