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

A second network of the same architecture is the learned preconditioner
(``Model.lp_pcg``): it maps features of a residual r over a node's neighbourhood (the
Jacobi-scaled residual of the node and of its neighbours one and two nodes away along each
axis, the local spacings, the face flag, the magnetic fraction, the coordinates) to the
error e = K^-1 r at the node, and preconditions a flexible CG whose stop test is the Jacobi
PCG's.  It is trained on (r, e) pairs from the same plain loops: at each field step, the
residuals after j Jacobi-PCG iterations from zero and their errors against the converged
potential.

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


def lp_features(model, r, X=None):
    """bf16 residual-feature rows (as int16) of every node for the residual r (device f64[N])."""
    torch = _torch()
    X = torch.empty((model.N, N_IN), dtype=torch.int16, device=r.device) if X is None else X
    model._sync()
    _lib.check(model._lib.mm_lp_features(model._ctx, _ptr(r, model.N), _ptr(X, model.N * N_IN, "int16")),
               "mm_lp_features")
    return X


def collect_both(model, params, steps, hstep, nodes_ws, nodes_lp, lp_iters=(0, 8, 32, 128, 512), seed=0,
                 log=None):
    """Training samples of both networks along the plain FP64 loop of ``model`` (``steps``
    field steps of ``hstep`` from hstart): per converged field step, (features of m, u) at
    ``nodes_ws`` random nodes for the warm start, and for each j in ``lp_iters`` the
    residual r_j = b - K x_j after j Jacobi-PCG iterations from zero (b = D m) with its
    error e_j = u - x_j: (features of r_j, e_j / s_j) at ``nodes_lp`` random nodes, s_j the
    RMS of r_j / diag(K).  Returns (X_ws, u, X_lp, t_lp)."""
    torch = _torch()
    from . import _lib as L
    dev = f"cuda:{model.device}"
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    model.lp_setup()
    dex, dk = np.empty(3 * model.N), np.empty(model.N)
    _lib.check(model._lib.mm_get_diag_host(model._ctx, dex.ctypes.data_as(_lib.c_double_p),
                                           dk.ctypes.data_as(_lib.c_double_p)), "diag")
    invd = torch.tensor(1.0 / dk, dtype=torch.float64, device=dev)
    m = model.to_device(model.uniform_m(params["m"]))
    m_ref, u, gF = m.clone(), model.empty1(), model.empty3()
    X = torch.empty((model.N, N_IN), dtype=torch.int16, device=dev)
    Xw, uw, Xl, tl = [], [], [], []
    hext = params["hstart"]
    st = L.Stats()
    for k in range(steps):
        model.minimize(m, hext, params["h"], params, m_ref, u, gF, st)
        features(model, m, X)
        idx = torch.randint(0, model.N, (nodes_ws,), generator=gen, device=dev)
        Xw.append(X[idx].clone())
        uw.append(u[idx].float())
        b = model.divergence(m)
        for j in lp_iters:
            xj = model.empty1()
            if j:
                model.pcg(b, xj, 1e-30, maxiter=j)
            r = b - model.apply_K(xj)
            e = u - xj
            sj = float(torch.sqrt(torch.mean((r * invd) ** 2)))
            if not sj > 0:
                continue
            lp_features(model, r, X)
            idx = torch.randint(0, model.N, (nodes_lp,), generator=gen, device=dev)
            Xl.append(X[idx].clone())
            tl.append((e[idx] / sj).float())
        if log:
            log(f"[collect] step {k}: h = {hext:.2f} T, M.h = {model.compute_mh(m, params['h']):.4f}")
        hext += hstep
    return torch.cat(Xw), torch.cat(uw), torch.cat(Xl), torch.cat(tl)


def collect_dense(config, seed, params, steps, hstep, nodes, log=None, targets="fp64"):
    """(X, u) warm-start samples of one microstructure at every one of ``steps`` field
    steps of ``hstep`` from hstart: the states m come from a sweep of the mixed form
    (FP32 inner solve with the multigrid preconditioner, FP64 outer residual; its field
    steps agree with plain FP64 to ~1e-8), and each target u is the potential of that m
    at ``nodes`` random nodes: targets "fp64", the plain FP64 Jacobi PCG's
    (scripts/solvers.py:221-231 at the loop's tolerance, started from the previous target);
    "mixed", the sweep's own last potential solve (the same FP64 residual tolerance).
    Returns (X, u)."""
    torch = _torch()
    from . import _lib as L
    from .generate import build_config
    from .model import Model
    g = build_config(config, seed=seed)
    mix = Model.from_structured(g, precision="mixed")
    mix.multigrid(8)
    f64 = Model.from_structured(g) if targets == "fp64" else mix
    dev = f"cuda:{mix.device}"
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    m = mix.to_device(mix.uniform_m(params["m"]))
    m_ref, u, gF, st = m.clone(), mix.empty1(), mix.empty3(), L.Stats()
    u64 = f64.empty1()
    X = torch.empty((f64.N, N_IN), dtype=torch.int16, device=dev)
    Xs, us = [], []
    hext = params["hstart"]
    for k in range(steps):
        mix.minimize(m, hext, params["h"], params, m_ref, u, gF, st)
        if targets == "fp64":
            b = f64.divergence(m)
            f64.pcg(b, u64, params["tol_u"])
        else:
            u64 = u
        features(f64, m, X)
        idx = torch.randint(0, f64.N, (nodes,), generator=gen, device=dev)
        Xs.append(X[idx].clone())
        us.append(u64[idx].float())
        if log and (k % 100 == 0 or k == steps - 1):
            log(f"[collect] seed {seed} step {k}: h = {hext:.3f} T, M.h = {mix.compute_mh(m, params['h']):.4f}")
        hext += hstep
    mix.close()
    if f64 is not mix:
        f64.close()
    return torch.cat(Xs), torch.cat(us)


def train_dense(config, seeds, steps, hstep, nodes, train_steps, batch, lr, out,
                log=lambda msg: print(msg, flush=True), budget_s=None, targets="fp64"):
    """The warm-start network on dense snapshots (collect_dense) of the microstructures
    ``seeds``: len(seeds) * steps (m, u) pairs; with ``budget_s``, the collection stops
    before a seed that would end past that many seconds (at the mean time per seed so far),
    and the saved provenance lists the seeds used."""
    torch = _torch()
    from .generate import build_config, loop_params
    from .model import Model
    p = loop_params(config)
    parts = []
    t0 = time.perf_counter()
    used = []
    for sd in seeds:
        el = time.perf_counter() - t0
        if budget_s is not None and used and el + el / len(used) > budget_s:
            log(f"[learned] time budget: {len(used)} of {len(seeds)} seeds collected ({el:.1f} s)")
            break
        parts.append(collect_dense(config, sd, p, steps, hstep, nodes,
                                   log=lambda msg: log(f"{msg} ({time.perf_counter() - t0:.1f} s)"), targets=targets))
        used.append(sd)
        log(f"[learned] seed {sd}: {parts[-1][0].shape[0]} samples ({time.perf_counter() - t0:.1f} s)")
    seeds = used
    X = torch.cat([q[0] for q in parts])
    u = torch.cat([q[1] for q in parts])
    del parts
    model = Model.from_structured(build_config(config, seed=seeds[0]))
    net = Network(model, max_batch=batch, seed=0)
    hist = train(net, X, u, steps=train_steps, batch=batch, lr=lr, log=log, log_every=max(train_steps // 10, 1))
    net.save(out, config=config, seeds=np.asarray(seeds), steps=steps, hstep=hstep, nodes_per_step=nodes,
             snapshots=np.int64(len(seeds) * steps), samples=np.int64(X.shape[0]), states="mixed-form sweep",
             targets="plain FP64 Jacobi PCG" if targets == "fp64" else "mixed form, FP64 residual tolerance",
             train_steps=train_steps, batch=batch, lr=lr,
             loss_first=hist[0][1], loss_last=hist[-1][1])
    log(f"[learned] warm start: {X.shape[0]} samples from {len(seeds) * steps} snapshots, loss {hist[0][1]:.3e} -> "
        f"{hist[-1][1]:.3e}, saved {out} ({time.perf_counter() - t0:.1f} s)")
    net.close()
    model.close()
    return hist


def train(net, X, u, steps=2000, batch=16384, lr=1e-3, lr_min=1e-5, seed=0, log=None, log_every=200,
          scale=None):
    """Adam on mean((net(X) - u / scale)^2) with a cosine learning-rate schedule;
    ``net.scale`` is set to ``scale`` or else the standard deviation of u.  Returns the
    loss history."""
    torch = _torch()
    net.scale = float(scale) if scale is not None else (float(u.double().std().item()) or 1.0)
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


def train_both(config, seeds, steps, hstep, nodes_ws, nodes_lp, train_steps, batch, lr, out_ws, out_lp,
               log=lambda msg: print(msg, flush=True)):
    """Both networks for a generated workload from the plain FP64 loops of the
    microstructures ``seeds``: the warm start on (m, u) snapshots, the preconditioner on
    (r, e) snapshots (its targets are already in units of the residual scale)."""
    torch = _torch()
    from .generate import build_config, loop_params
    from .model import Model
    p = loop_params(config)
    parts = []
    t0 = time.perf_counter()
    for sd in seeds:
        model = Model.from_structured(build_config(config, seed=sd))
        parts.append(collect_both(model, p, steps, hstep, nodes_ws, nodes_lp, seed=sd,
                                  log=lambda msg: log(f"{msg} (seed {sd}, {time.perf_counter() - t0:.1f} s)")
                                  if msg.startswith("[collect] step 0") else None))
        model.close()
        log(f"[learned] seed {sd}: {parts[-1][0].shape[0]} warm-start and {parts[-1][2].shape[0]} "
            f"preconditioner samples ({time.perf_counter() - t0:.1f} s)")
    Xw = torch.cat([q[0] for q in parts])
    uw = torch.cat([q[1] for q in parts])
    Xl = torch.cat([q[2] for q in parts])
    tl = torch.cat([q[3] for q in parts])
    del parts
    model = Model.from_structured(build_config(config, seed=seeds[0]))
    snaps = dict(config=config, seeds=np.asarray(seeds), steps=steps, hstep=hstep, train_steps=train_steps,
                 batch=batch, lr=lr)
    hists = []
    for name, X, t, out, scale in (("warm start", Xw, uw, out_ws, None), ("preconditioner", Xl, tl, out_lp, None)):
        net = Network(model, max_batch=batch, seed=0)
        hist = train(net, X, t, steps=train_steps, batch=batch, lr=lr, log=log,
                     log_every=max(train_steps // 10, 1), scale=scale)
        net.save(out, samples=np.int64(X.shape[0]), snapshots=np.int64(len(seeds) * steps *
                 (1 if name == "warm start" else 5)), loss_first=hist[0][1], loss_last=hist[-1][1], **snaps)
        log(f"[learned] {name}: {X.shape[0]} samples, loss {hist[0][1]:.3e} -> {hist[-1][1]:.3e}, saved {out} "
            f"({time.perf_counter() - t0:.1f} s)")
        net.close()
        hists.append(hist)
    model.close()
    return hists


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
    ap.add_argument("--both", action="store_true",
                    help="train the warm start and the learned preconditioner together (data/warmstart_*.npz, "
                         "data/precond_*.npz)")
    ap.add_argument("--nodes-lp", type=int, default=4096, help="--both: sampled nodes per residual snapshot")
    ap.add_argument("--out-dir", default=None, help="--both: directory of the two .npz files (default: the package's)")
    ap.add_argument("--budget", type=float, default=None, help="--dense: seconds of data collection at most")
    ap.add_argument("--targets", default="fp64", choices=["fp64", "mixed"],
                    help="--dense: the potential of each snapshot from plain FP64 or from the sweep's mixed form")
    ap.add_argument("--dense", action="store_true",
                    help="the warm start only, on a snapshot at every field step of mixed-form sweeps with plain "
                         "FP64 targets (collect_dense)")
    a = ap.parse_args()
    if a.dense:
        out = a.out or str(pathlib.Path(__file__).resolve().parent / "data" / f"warmstart_{a.config}.npz")
        train_dense(a.config, a.seeds, a.steps, a.hstep, a.nodes, a.train_steps, a.batch, a.lr, out,
                    budget_s=a.budget, targets=a.targets)
        raise SystemExit(0)
    if a.both:
        d = pathlib.Path(a.out_dir) if a.out_dir else pathlib.Path(__file__).resolve().parent / "data"
        d.mkdir(exist_ok=True)
        train_both(a.config, a.seeds, a.steps, a.hstep, a.nodes, a.nodes_lp, a.train_steps, a.batch, a.lr,
                   str(d / f"warmstart_{a.config}.npz"), str(d / f"precond_{a.config}.npz"))
        raise SystemExit(0)
    out = a.out or str(pathlib.Path(__file__).resolve().parent / "data" / f"warmstart_{a.config}.npz")
    pathlib.Path(out).parent.mkdir(exist_ok=True)
    train_for_config(a.config, a.seeds, a.steps, a.hstep, a.nodes, a.train_steps, a.batch, a.lr, out)
