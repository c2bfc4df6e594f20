"""Benchmark: hysteresis field steps per second on one MI355X.

One step = one field value of the hysteresis sweep: the full nonlinear-CG
minimisation (``minimize``, loop.py:58-65 of the reference) at that field, on the
synthetic 64-grain Voronoi cube of 125^3 voxels (~2.0 M magnet nodes, 3.65 M nodes,
21.5 M Kuhn tetrahedra) from +8 T in -0.05 T steps.  The first W field steps are
warm-up; the next K are timed.  Inputs are resident in HBM before timing starts.

The value is the throughput of the form chosen by --precision (default: FP32 inner
solve with FP64 outer refinement, FP64 dots / energies / line search); the plain FP64
path then runs the same steps and the line carries its throughput, the speedup and the
form's errors against it (M.h, energies, potential u, demag field G u).

Prints one JSON line.  The FP64 oracle is timed on the host cores as a reported
baseline in a child process, concurrently with the GPU runs.
"""
import argparse
import json
import os
import pathlib
import subprocess
import sys
import tempfile
import time

ROOT = pathlib.Path(__file__).resolve().parent
for _p in (ROOT, ROOT / "mammos-mumag_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

METRIC = "hysteresis field steps/s @2M nodes: learned+FP32 vs plain FP64 on one MI355X"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the two PCG kernels and of the level-0 multigrid smoothing step
# on the voronoi64_250nm grid (structured path), from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (median KiB per dispatch) scaled by the same counters on the 512 MiB device
# copy and fill of --traffic-probe (FETCH_SIZE reads half the copied bytes: x 2;
# WRITE_SIZE the filled bytes: x 1); see
# profiles/pcg_kernels_{fp64,mixed_jacobi}_voronoi125.md and profiles/mg_vcycle_mixed_voronoi125.md.
TRAFFIC = {
    ("fp64", "spmv"): 127833702, ("fp64", "update"): 233822003,
    ("mixed", "spmv"): 66500096, ("mixed", "update"): 116983808,
    ("mixed", "mg_smooth"): 45730406,
}


def key_of(k):
    return k["key"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# --------------------------------------------------------------------------- CPU leg
def cpu_worker(cfg, out_path, seed, full_step=False):
    """Unit costs of the FP64 oracle on this host, at the benchmark's own mesh."""
    import numpy as np

    from mammos_mumag_amd.generate import build_config, loop_params
    from oracle.assemble import Model
    from oracle.energies import total_eg
    from oracle.loop import Minimizer
    from oracle.solvers import jax_cg

    t0 = time.perf_counter()
    g = build_config(cfg, seed=seed)
    M = Model(g.coords(), g.tets(), g.elem_tags(), g.krn())
    t_setup = time.perf_counter() - t0
    p = loop_params(cfg)
    m = M.uniform_m(p["m"])
    b = M.Dxyz[0] @ m[0::3] + M.Dxyz[1] @ m[1::3] + M.Dxyz[2] @ m[2::3]
    n_pcg = 100
    t0 = time.perf_counter()
    u, it = jax_cg(M.K, M.diagK, b, np.zeros(M.N), tol=1e-30, maxiter=n_pcg)
    t_pcg = (time.perf_counter() - t0) / it
    # one evaluation without PCG iterations (tol 1: the start already satisfies it)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        E, gp, _, gF = total_eg(M, m, 1.0, p["hstart"], p["h"], 1, u)
    t_eval = (time.perf_counter() - t0) / reps
    mini = Minimizer(M, p, m)
    mini.gF = gF
    t0 = time.perf_counter()
    mini.M_inv(gp)
    n_tcg = mini.stats["tcg_iters"]
    t_tcg = (time.perf_counter() - t0) / max(n_tcg, 1)
    sample_s = t_pcg * it + t_eval * reps + t_tcg * n_tcg
    threads = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    res = dict(t_pcg_iter=t_pcg, t_eval=t_eval, t_tcg_iter=t_tcg, setup_s=t_setup,
               sample_s=sample_s, n_pcg=it, n_tcg=n_tcg, cores=1,
               host_cores=len(os.sched_getaffinity(0)), threads_env=threads)
    if full_step:
        # one complete field step of the oracle (loop.py:58-65 at hstart from m0 = +z, u = 0),
        # against the unit-cost model applied to its own counts
        mini = Minimizer(M, p, m)
        t0 = time.perf_counter()
        mini.minimize(m.copy(), p["hstart"])
        full_s = time.perf_counter() - t0
        sc = mini.stats
        model_s = sc["pcg_iters"] * t_pcg + sc["calls"] * t_eval + sc["tcg_iters"] * t_tcg
        res.update(full_step_s=full_s, full_step_counts={k: int(v) for k, v in sc.items()},
                   unit_cost_model_s=model_s, model_over_measured=model_s / full_s)
    with open(out_path, "w") as f:
        json.dump(res, f)


def start_cpu_leg(cfg, seed):
    fd, path = tempfile.mkstemp(suffix=".json", prefix="mumag_cpu_")
    os.close(fd)
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    proc = subprocess.Popen([sys.executable, str(ROOT / "bench.py"), "--cpu-worker", path,
                             "--config", cfg, "--seed", str(seed)], env=env)
    return proc, path


# --------------------------------------------------------------------------- GPU leg
def gpu_run(args, precision, traffic_probe=False, learned=None, multigrid=0, amg=False):
    """Warm-up and timed field steps of one form; returns throughput, counters, kernel
    timing and the per-step records (M.h, E) plus the final potential and demag field."""
    import numpy as np
    import torch

    from mammos_mumag_amd import Model, _lib
    from mammos_mumag_amd.generate import build_config, loop_params

    t0 = time.perf_counter()
    g = build_config(args.config, seed=args.seed)
    n_nodes, n_elems = len(g.xs) * len(g.ys) * len(g.zs), 6 * (len(g.xs) - 1) * (len(g.ys) - 1) * (len(g.zs) - 1)
    log(f"[bench] {precision}: mesh {args.config}: {n_nodes} nodes, {n_elems} tets ({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    model = Model.from_structured(g, structured=not args.general, precision=precision,
                                  inner_tol=args.inner_tol, eval_precision=args.eval_precision)
    levels = []
    if multigrid:
        model.multigrid(multigrid)
        levels = model.mg_levels()
    amg_sizes, amg_setup_s = [], None
    if amg:
        t1 = time.perf_counter()
        model.amg()
        amg_setup_s = time.perf_counter() - t1
        amg_sizes = model.amg_levels()
        log(f"[bench] AMG set-up: levels {amg_sizes} ({amg_setup_s:.2f} s)")
    if (multigrid or amg) and not args.no_secant:
        model.secant_start(True)
    tag = (precision + (f" + multigrid ({len(levels)} levels)" if multigrid else "")
           + (f" + AMG ({len(amg_sizes)} levels)" if amg else "") + (" + learned" if learned else "")
           + (" + secant" if (multigrid or amg) and not args.no_secant else ""))
    if learned:
        from mammos_mumag_amd.learned import Network
        net = Network(model, max_batch=32, seed=0)
        net.load(learned)
        model.warm_start(net)
        model.warm_start_policy(args.ws_policy)
    log(f"[bench] {tag}: context {model.info} ({time.perf_counter() - t0:.1f} s)")
    p = loop_params(args.config)
    m = model.to_device(model.uniform_m(p["m"]))
    m_ref = m.clone()
    u = model.empty1()
    gF = model.empty3()
    hext = p["hstart"]
    st = _lib.Stats()
    for w in range(args.warmup):
        t1 = time.perf_counter()
        E = model.minimize(m, hext, p["h"], p, m_ref, u, gF, st)
        log(f"[bench] {tag}: warmup step {w}: h={hext:.2f} T E={E:.6e} {time.perf_counter() - t1:.3f} s "
            f"{st.as_dict()}")
        hext += p["hstep"]
    st = _lib.Stats()
    model.timing(True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    energies = []
    for k in range(args.steps):
        energies.append(model.minimize(m, hext, p["h"], p, m_ref, u, gF, st))
        hext += p["hstep"]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    timing = model.timing_get()
    model.timing(False)
    mh = model.compute_mh(m, p["h"])
    log(f"[bench] {tag}: timed {args.steps} steps: {elapsed:.3f} s, last h={hext - p['hstep']:.2f} T, "
        f"M.h={mh:.6f} T, E={energies[-1]:.6e}, {st.as_dict()}, {timing}")
    if traffic_probe:
        # reference dispatches for the HBM byte counters: a 512 MiB fill (torch.ones,
        # writes 524288 KiB) and a 512 MiB device copy (reads 524288 KiB)
        src = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] traffic probe: copied 512 MiB")
    ws = dict(model.warm_start_stats(), **model.warm_start_report()) if learned else None
    res = dict(elapsed=elapsed, stats=st.as_dict(), timing=timing, info=model.info, warm_start=ws,
               mg_levels=" -> ".join("x".join(map(str, d)) for d in levels), levels=levels,
               amg_levels=amg_sizes, amg_setup_s=amg_setup_s,
               n_voxels=int(CONFIGS_N(args.config)) ** 3,
               n_nodes=n_nodes, n_elems=n_elems, n_mag_nodes=int(np.sum(model.meas > 0)),
               mh=mh, energies=energies, u=u.cpu().numpy(), hd=model.gradient(u).cpu().numpy())
    model.close()
    del m, m_ref, u, gF, model
    torch.cuda.empty_cache()
    return res


def loop_run(args, precision, multigrid=0, learned=None, perturb=0.0):
    """The field sweep of --config from hstart until ``--loop-after`` steps past the reversal
    (the first step with M.h < 0), as scripts/loop.py:89-124 runs it with mfinal -2; per
    step the field, M.h, E, wall seconds and the counters of that step."""
    import torch

    from mammos_mumag_amd import Model, _lib
    from mammos_mumag_amd.generate import build_config, loop_params

    g = build_config(args.config, seed=args.seed)
    model = Model.from_structured(g, precision=precision, eval_precision=args.eval_precision)
    if multigrid:
        model.multigrid(multigrid)
        if not args.no_secant:
            model.secant_start(True)
    tag = (precision + (" + multigrid" if multigrid else "") + (" + secant" if multigrid and not args.no_secant else "")
           + (" + learned" if learned else ""))
    if precision == "mixed":
        tag += f" (evaluation {args.eval_precision})"
    if learned:
        from mammos_mumag_amd.learned import Network
        net = Network(model, max_batch=32, seed=0)
        net.load(learned)
        model.warm_start(net)
        model.warm_start_policy(args.ws_policy)
    p = loop_params(args.config)
    m0 = model.uniform_m(p["m"])
    m_ref = model.to_device(m0)
    if perturb:
        import numpy as np
        m0 = m0 * (1 + perturb * np.random.default_rng(1).normal(size=m0.shape))
        tag += f", initial m * (1 + {perturb:g} N(0,1))"
    m = model.to_device(m0)
    u, gF = model.empty1(), model.empty3()
    st = _lib.Stats()
    n_steps = int(abs(p["hfinal"] - p["hstart"]) / abs(p["hstep"])) + 1
    hext, rec, prev, after = p["hstart"], [], st.as_dict(), None
    t_loop = time.perf_counter()
    for k in range(n_steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E = model.minimize(m, hext, p["h"], p, m_ref, u, gF, st)
        mh = model.compute_mh(m, p["h"])
        dt = time.perf_counter() - t0
        cur = st.as_dict()
        d = {key: cur[key] - prev[key] for key in cur}
        prev = cur
        rec.append(dict(h=hext, mh=mh, E=E, s=dt, **d))
        if k % 20 == 0 or mh < 0 or d["ncg_iters"] > 50:
            log(f"[loop] {tag}: step {k} h={hext:+.2f} T M.h={mh:+.6f} T {dt:.3f} s "
                f"ncg {d['ncg_iters']} pcg {d['pcg_iters']} backtracks {d['backtracks']}")
        if mh < 0 and after is None:
            after = k
        if after is not None and k - after >= args.loop_after:
            break
        hext += p["hstep"]
    total = time.perf_counter() - t_loop
    V, moment = model.V, float((model.meas).sum())
    model.close()
    del m, m_ref, u, gF, model
    torch.cuda.empty_cache()
    return dict(tag=tag, records=rec, seconds=total, js_sat=moment / V)


def loop_main(args, learned):
    """--loop: the chosen form and plain FP64 through the reversal; one JSON line."""
    import numpy as np
    forms = [("mixed", args.multigrid, learned), ("fp64", 0, None)]
    if args.loop_perturb:
        forms.append(("fp64", 0, None, args.loop_perturb))
    if args.loop_fp64_only:
        forms = forms[1:]
    runs = [loop_run(args, *f) for f in forms]
    out = {"loop": args.config, "forms": []}
    ref = runs[0 if args.loop_fp64_only else 1]
    n = min(len(r["records"]) for r in runs)
    sat = args.loop_window
    for r in runs:
        rec = r["records"]
        mh = np.array([x["mh"] for x in rec])
        sw = int(np.argmax(mh < 0)) if np.any(mh < 0) else None
        t = np.array([x["s"] for x in rec])
        f = dict(form=r["tag"], steps=len(rec), reversal_step=sw,
                 reversal_field_T=None if sw is None else rec[sw]["h"],
                 steps_per_s_whole=round(len(rec) / float(t.sum()), 4),
                 steps_per_s_saturated=round(sat / float(t[:sat].sum()), 4),
                 slowest_step_s=round(float(t.max()), 3),
                 ncg_per_step=round(float(np.mean([x["ncg_iters"] for x in rec])), 2),
                 pcg_per_step=round(float(np.mean([x["pcg_iters"] for x in rec])), 2),
                 backtracks=int(sum(x["backtracks"] for x in rec)),
                 restarts=int(sum(x["restarts_g"] + x["restarts_z"] for x in rec)))
        if r is not ref:
            a = np.array([x["mh"] for x in rec[:n]])
            b = np.array([x["mh"] for x in ref["records"][:n]])
            ea = np.array([x["E"] for x in rec[:n]])
            eb = np.array([x["E"] for x in ref["records"][:n]])
            f["max_abs_dM_over_Ms"] = float(np.max(np.abs(a - b)) / ref["js_sat"])
            f["max_dE_over_maxE"] = float(np.max(np.abs(ea - eb)) / np.max(np.abs(eb)))   # E crosses zero
            f["speedup_whole"] = round(float(sum(x["s"] for x in ref["records"][:n]) / sum(x["s"] for x in rec[:n])), 3)
        out["forms"].append(f)
    if args.loop_out:
        with open(args.loop_out, "w") as fh:
            json.dump(dict(out, records={r["tag"]: r["records"] for r in runs}), fh)
    print(json.dumps(out), flush=True)
    return 0


def avalanche_main(args, learned):
    """--avalanche K: plain FP64 runs the sweep to step K - 1; from that state, step K (the
    reversal avalanche) is minimised again by plain FP64 from the exact state and from states
    perturbed by 1e-15 relative (--avalanche-seeds), and by the chosen form.  Shows how far
    rounding alone moves the end state of an avalanche in the FP64 method itself."""
    import numpy as np
    import torch

    from mammos_mumag_amd import Model, _lib
    from mammos_mumag_amd.generate import build_config, loop_params

    g = build_config(args.config, seed=args.seed)
    p = loop_params(args.config)
    f64 = Model.from_structured(g)
    m = f64.to_device(f64.uniform_m(p["m"]))
    m_ref, u, gF, st = m.clone(), f64.empty1(), f64.empty3(), _lib.Stats()
    hext = p["hstart"]
    t0 = time.perf_counter()
    for k in range(args.avalanche):
        f64.minimize(m, hext, p["h"], p, m_ref, u, gF, st)
        if k % 20 == 0 or k >= args.avalanche - 15:
            log(f"[avalanche] fp64 step {k} h={hext:+.2f} T M.h={f64.compute_mh(m, p['h']):+.6f} T "
                f"({time.perf_counter() - t0:.0f} s)")
        hext += p["hstep"]
    mh0 = f64.compute_mh(m, p["h"])
    saved = (m.clone(), u.clone(), gF.clone())
    out = dict(avalanche=args.config, step=args.avalanche, h=hext, mh_before=mh0, runs=[])

    def one(model, tag, mm, uu, gg):
        s2 = _lib.Stats()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        E = model.minimize(mm, hext, p["h"], p, m_ref, uu, gg, s2)
        dt = time.perf_counter() - t1
        r = dict(run=tag, mh=model.compute_mh(mm, p["h"]), E=E, seconds=round(dt, 3),
                 ncg_iters=s2.ncg_iters, pcg_iters=s2.pcg_iters, backtracks=s2.backtracks)
        log(f"[avalanche] {r}")
        out["runs"].append(r)

    mag = torch.tensor(np.repeat(f64.meas > 0, 3), dtype=torch.float64, device="cuda")
    runs = [(None, 0.0)] + [(sd, float(e)) for e in args.avalanche_eps.split(",")
                            for sd in range(1, args.avalanche_seeds + 1)]
    for seed, eps in runs:
        mm, uu, gg = (x.clone() for x in saved)
        if seed is not None:
            gen = torch.Generator(device="cuda").manual_seed(seed)
            mm *= 1 + eps * torch.randn(mm.shape, dtype=torch.float64, device="cuda", generator=gen) * mag
        one(f64, "fp64" if seed is None else f"fp64, m * (1 + {eps:g} N(0,1)) seed {seed}", mm, uu, gg)
    f64.close()
    torch.cuda.empty_cache()
    mix = Model.from_structured(g, precision="mixed", eval_precision=args.eval_precision)
    mix.multigrid(args.multigrid)
    if not args.no_secant:
        mix.secant_start(True)
    if learned:
        from mammos_mumag_amd.learned import Network
        net = Network(mix, max_batch=32, seed=0)
        net.load(learned)
        mix.warm_start(net)
        mix.warm_start_policy(args.ws_policy)
    mm, uu, gg = (x.clone() for x in saved)
    one(mix, "chosen form (FP32 inner solve, multigrid" + (", secant start" if not args.no_secant else "")
        + (", learned warm start" if learned else "") + ")", mm, uu, gg)
    mix.close()
    print(json.dumps(out), flush=True)
    return 0


def reversal_check_main(args):
    """--reversal-check: the chosen form (FP32 inner solve with multigrid, no network) sweeps
    the field of --config from --rc-hstart by --rc-hstep until the first step with M.h < 0
    (the reversal, step K) or the --rc-time-limit; the states after the last steps are kept
    on the device and plain FP64 then re-solves step K - 1 from the chosen form's state after
    K - 2 and step K from its state after K - 1 (--rc-fp64): the switching step of FP64 from
    the same history, and |dM|/Ms at those steps.  One JSON line."""
    import numpy as np
    import torch

    from mammos_mumag_amd import Model, _lib
    from mammos_mumag_amd.generate import build_config, loop_params

    g = build_config(args.config, seed=args.seed)
    p = loop_params(args.config)
    h0 = p["hstart"] if args.rc_hstart is None else args.rc_hstart
    dh = p["hstep"] if args.rc_hstep is None else args.rc_hstep
    t_set = time.perf_counter()
    mix = Model.from_structured(g, precision="mixed", eval_precision=args.eval_precision)
    mix.multigrid(args.multigrid)
    if not args.no_secant:
        mix.secant_start(True)
    log(f"[reversal] {args.config}: {mix.info} multigrid {mix.mg_levels()} ({time.perf_counter() - t_set:.1f} s)")
    m = mix.to_device(mix.uniform_m(p["m"]))
    m_ref, u, gF, st = m.clone(), mix.empty1(), mix.empty3(), _lib.Stats()
    ms_sat = float(mix.meas.sum()) / mix.V
    keep = []   # (k, h, m, u, gF) before each of the last steps, or before the --rc-resolve steps
    rec, K = [], None
    hext, k = h0, 0
    t_loop = time.perf_counter()
    while True:
        if args.rc_resolve is None or k in args.rc_resolve:
            keep.append((k, hext, m.clone(), u.clone(), gF.clone()))
        if args.rc_resolve is None:
            keep = keep[-2:]
        prev = st.as_dict()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E = mix.minimize(m, hext, p["h"], p, m_ref, u, gF, st)
        mh = mix.compute_mh(m, p["h"])
        dt = time.perf_counter() - t0
        cur = st.as_dict()
        d = {key: cur[key] - prev[key] for key in cur}
        rec.append(dict(step=k, h=round(hext, 6), mh=mh, E=E, s=round(dt, 3), **d))
        log(f"[reversal] chosen form step {k} h={hext:+.3f} T M.h={mh:+.6f} T {dt:.2f} s ncg {d['ncg_iters']} "
            f"pcg {d['pcg_iters']} backtracks {d['backtracks']}")
        if mh < 0:
            K = k
            break
        if time.perf_counter() - t_loop > args.rc_time_limit or (args.rc_stop_step is not None and k >= args.rc_stop_step):
            break
        hext += dh if args.rc_fine_from is None or hext + dh > args.rc_fine_from else args.rc_hstep_fine
        k += 1
    loop_s = time.perf_counter() - t_loop
    out = dict(reversal_check=args.config, hstart=h0, hstep=dh, fine_from=args.rc_fine_from,
               hstep_fine=args.rc_hstep_fine if args.rc_fine_from is not None else None, js_sat=ms_sat, steps=len(rec),
               reversal_step=K, reversal_field_T=None if K is None else rec[K]["h"], loop_seconds=round(loop_s, 1),
               chosen=dict(mean_s_per_step=round(loop_s / len(rec), 3),
                           ncg_per_step=round(float(np.mean([r["ncg_iters"] for r in rec])), 2),
                           pcg_per_step=round(float(np.mean([r["pcg_iters"] for r in rec])), 2),
                           slowest_step_s=max(r["s"] for r in rec)),
               records=rec)
    mix.close()
    del mix
    torch.cuda.empty_cache()
    print(json.dumps(dict(out, partial=True)), flush=True)
    if args.rc_fp64 and (args.rc_resolve is not None or (K is not None and len(keep) == 2)):
        f64 = Model.from_structured(g)
        cmp_ = []
        for (kk, hh, m0, u0, g0) in keep:   # steps K - 1 and K (or the --rc-resolve steps)
            st2 = _lib.Stats()
            mm, uu, gg = m0.clone(), u0.clone(), g0.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            E2 = f64.minimize(mm, hh, p["h"], p, m_ref, uu, gg, st2)
            dt = time.perf_counter() - t0
            mh2 = f64.compute_mh(mm, p["h"])
            r = rec[kk]
            c = dict(step=kk, h=hh, mh_fp64=mh2, mh_chosen=r["mh"], dM_over_Ms=abs(mh2 - r["mh"]) / ms_sat,
                     E_fp64=E2, E_chosen=r["E"], seconds_fp64=round(dt, 1), seconds_chosen=r["s"],
                     fp64_counts=st2.as_dict())
            log(f"[reversal] fp64 re-solve {c}")
            print(json.dumps(dict(partial=True, fp64_resolve=c)), flush=True)
            cmp_.append(c)
        out["fp64_resolves"] = cmp_
        if args.rc_resolve is None:
            out["fp64_switches_at_same_step"] = bool(cmp_[0]["mh_fp64"] >= 0 and cmp_[1]["mh_fp64"] < 0)
        f64.close()
    print(json.dumps(out), flush=True)
    return 0


def precond_eval_main(args):
    """--precond-eval: the learned preconditioner on the benchmark geometry with a
    microstructure no network saw (--seed): potential solves b = D m of converged field
    steps of plain FP64, from zero and from the previous step's potential, by the Jacobi
    PCG (fused production kernels, and the flexible-CG driver without the network), the
    flexible CG with the network every `every`-th iteration, and the chosen form's FP32
    multigrid solve.  One JSON line."""
    import numpy as np
    import torch

    from mammos_mumag_amd import Model, _lib
    from mammos_mumag_amd.generate import build_config, loop_params
    from mammos_mumag_amd.learned import Network

    g = build_config(args.config, seed=args.seed)
    p = loop_params(args.config)
    f64 = Model.from_structured(g)
    f64.lp_setup()
    net = Network(f64, max_batch=32, seed=0)
    meta = net.load(args.precond)
    log(f"[precond] network {args.precond}: trained on seeds {list(meta.get('seeds', []))}, "
        f"loss {float(meta.get('loss_first', 0)):.3e} -> {float(meta.get('loss_last', 0)):.3e}")
    mix = Model.from_structured(g, precision="mixed")
    mix.multigrid(args.multigrid)
    m = f64.to_device(f64.uniform_m(p["m"]))
    m_ref, u, gF, st = m.clone(), f64.empty1(), f64.empty3(), _lib.Stats()
    fields = [p["hstart"] + k * p["hstep"] for k in range(args.precond_steps)]
    rows = []
    u_prev = None

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    for k, hext in enumerate(fields):
        f64.minimize(m, hext, p["h"], p, m_ref, u, gF, st)
        b = f64.divergence(m)
        starts = [("zero", f64.empty1())] + ([("previous", u_prev)] if u_prev is not None else [])
        for sname, x0 in starts:
            res = dict(step=k, h=hext, start=sname)
            x = x0.clone()
            (it, _), t = timed(lambda: f64.pcg(b, x, 1e-10))
            res["jacobi_fused"] = dict(iters=it, ms=round(1e3 * t, 2))
            ref = x.clone()
            x = x0.clone()
            (it, _, _), t = timed(lambda: f64.lp_pcg(b, x, 1e-10, net=None))
            res["jacobi_fcg"] = dict(iters=it, ms=round(1e3 * t, 2),
                                     du=float(torch.max(torch.abs(x - ref)) / torch.max(torch.abs(ref))))
            for every in args.precond_every:
                x = x0.clone()
                (it, apps, rc), t = timed(lambda: f64.lp_pcg(b, x, 1e-10, net=net, every=every))
                res[f"learned_every{every}"] = dict(iters=it, apps=apps, ms=round(1e3 * t, 2), rc=rc,
                                                    du=float(torch.max(torch.abs(x - ref)) / torch.max(torch.abs(ref))))
            xm = x0.clone()
            (it, _), t = timed(lambda: mix.pcg(b, xm, 1e-10))
            res["multigrid_fp32"] = dict(iters=it, ms=round(1e3 * t, 2),
                                         du=float(torch.max(torch.abs(xm - ref)) / torch.max(torch.abs(ref))))
            log(f"[precond] {res}")
            rows.append(res)
        u_prev = u.clone()
    summary = {}
    for key in rows[0]:
        if key in ("step", "h", "start"):
            continue
        for sname in ("zero", "previous"):
            sel = [r[key] for r in rows if r["start"] == sname and key in r]
            if sel:
                summary[f"{key}/{sname}"] = dict(iters=round(float(np.mean([v["iters"] for v in sel])), 1),
                                                 ms=round(float(np.mean([v["ms"] for v in sel])), 2))
    out = dict(precond_eval=args.config, seed=args.seed, network=os.path.relpath(args.precond, ROOT),
               trained_on=[int(v) for v in np.ravel(meta.get("seeds", []))], steps=len(fields), summary=summary,
               rows=rows)
    print(json.dumps(out), flush=True)
    net.close()
    mix.close()
    f64.close()
    return 0


def roofline(res, precision, args):
    """Dominant PCG kernel of the timed steps: algorithmic bytes / measured launch time."""
    tm = res["timing"]
    prec = "FP32 inner solve" if precision == "mixed" else "FP64"
    kname = ("k_ka_grid (7-point Poisson stencil fused with p = z + beta p, p.q partials)"
             if res["info"]["structured"] else "k_ka_sell (SELL-64 K apply fused with p = z + beta p, p.q partials)")
    kernels = []
    # level-0 smoothing launches per PCG iteration: geometric V-cycle 6, AMG 3 (degree 2:
    # one step down, the first fused into the update, two up)
    smooth = (("mg_smooth", "k_amg_smooth<1> (Chebyshev smoothing step of the AMG V-cycle's finest level, "
                            "SELL-64 K), FP32 inner solve", 3) if res.get("amg_levels") else
              ("mg_smooth", "k_mg_stencil<SMOOTH> (Chebyshev smoothing step of the multigrid "
                            "V-cycle's finest level), FP32 inner solve", 6))
    # (timer, kernel, launches of that kernel per PCG iteration)
    for key, name, per_it in (("spmv", f"{kname}, {prec}", 1),
                              ("update", f"k_kb (x, r, z update, r.z and r.r partials), {prec}", 1),
                              smooth):
        if tm[f"{key}_launches"] > 0:
            avg_s = tm[f"{key}_ms"] / tm[f"{key}_launches"] / 1e3
            kernels.append(dict(key=key, kernel=name, total_ms=tm[f"{key}_ms"], avg_s=avg_s, per_it=per_it,
                                bytes=tm[f"{key}_bytes"], launches=tm[f"{key}_launches"]))
    if not kernels:
        return None
    # the dominant kernel: largest share of the PCG iteration (launch time x launches)
    k = max(kernels, key=lambda d: d["avg_s"] * d["per_it"])
    achieved = k["bytes"] / k["avg_s"] / 1e9
    traffic = (TRAFFIC.get((precision, key_of(k)))
               if args.config == "voronoi64_250nm" and res["info"]["structured"] else None)
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                kernel=k["kernel"], bytes_per_launch=k["bytes"], avg_launch_us=round(k["avg_s"] * 1e6, 2),
                sampled_launches=k["launches"],
                others=[dict(kernel=o["kernel"], avg_launch_us=round(o["avg_s"] * 1e6, 2),
                             achieved_GBs=round(o["bytes"] / o["avg_s"] / 1e9, 1))
                        for o in kernels if o is not k],
                per_pcg_iteration=k["per_it"])


def CONFIGS_N(cfg):
    from mammos_mumag_amd.generate import CONFIGS
    return CONFIGS[cfg]["n"]


def whole_step(res, precision, steps):
    """Algorithmic HBM bytes of the timed field steps, by kernel class, from the run's exact
    counts (PCG and truncated-CG iterations, energy evaluations, V-cycles that ran, true
    residual passes and corrections) and per-launch byte counts on the structured grid
    (N nodes, Nb magnet-block nodes, Nv voxels with 2-byte material ids); divided by the
    measured time per step.  Not counted: the NCG's reductions, retractions and line-search
    vector updates, finalize launches, copies and the network's inference."""
    if not res["info"]["structured"]:
        return None
    st, tm = res["stats"], res["timing"]
    N, Nb, Nv = res["n_nodes"], res["n_mag_nodes"], res["n_voxels"]
    s = 4 if precision == "mixed" else 8   # inner PCG / truncated-CG vectors
    cls = {}
    # KA: z, p_old read, p, q written; KB: x, r, p, q, diag (or the V-cycle weights) read,
    # x, r, z written
    cls["pcg_iterations"] = st["pcg_iters"] * 12 * s * N
    if precision == "mixed":
        lv = res["levels"]
        if lv and tm["vcycles"] >= 0:
            n0 = lv[0][0] * lv[0][1] * lv[0][2]
            coarse = sum(a * b * c for a, b, c in lv[1:])
            # level 0: 7 smoothing steps and the residual (b, x read, one vector written),
            # r read by the restriction, x read and written by the prolongation; coarser
            # levels: 3 smoothing passes (two of them two-step), residual, transfers
            cls["v_cycles"] = tm["vcycles"] * (108 * n0 + 72 * coarse)
        # true residual: x, b (FP64), diag (FP32) read, r, z (FP32) written; correction:
        # x (FP64), dx (FP32) read and written
        cls["true_residuals"] = tm["ru_passes"] * 28 * N + tm["corrections"] * 24 * N
    else:
        # per solve: K x0 (x read, y written) and the initial residual (b, Kx0, diag read;
        # r, z written)
        cls["pcg_starts"] = st["pcg_calls"] * (16 + 40) * N
    # truncated CG (magnet block): d' = -z + beta d, C d', H d' (z, d, m0, m0.gF read; d', Hd
    # written; voxel materials), the update (r, Hd, diag read; r, z written) and, once per
    # solve (one per NCG iteration), the solution p = sum alpha_k d'_k (each iteration's d'
    # read, p written)
    cls["truncated_cg"] = st["tcg_iters"] * ((34 * s) * Nb + 2 * Nv) + st["ncg_iters"] * 3 * s * Nb
    # evaluation (FP64): C m with D m (m read, gC and b written), G u (u read, g written),
    # finish (meas, m, gC, gd read; g_full, g_proj written)
    cls["evaluations"] = st["fcalls"] * (216 * Nb + 4 * Nv)
    total = sum(cls.values())
    t_step = res["elapsed"] / steps
    bps = total / steps
    return dict(bytes_per_step=int(bps), GBs=round(bps / t_step / 1e9, 1), frac=round(bps / t_step / 1e9 / HBM_PEAK_GBS, 4),
                by_class_MB_per_step={k: round(v / steps / 1e6, 1) for k, v in cls.items()},
                counts_per_step=dict(vcycles=round(tm["vcycles"] / steps, 2) if tm["vcycles"] >= 0 else None,
                                     ru_passes=round(tm["ru_passes"] / steps, 2),
                                     corrections=round(tm["corrections"] / steps, 2)))


def form_errors(r, ref):
    """Errors of a run against the FP64 run over the same field steps."""
    import numpy as np
    e = np.asarray(r["energies"])
    e0 = np.asarray(ref["energies"])
    return dict(mh_abs=float(abs(r["mh"] - ref["mh"])),
                E_rel_max=float(np.max(np.abs(e / e0 - 1))),
                u_rel=float(np.max(np.abs(r["u"] - ref["u"])) / np.max(np.abs(ref["u"]))),
                hdemag_rel=float(np.max(np.abs(r["hd"] - ref["hd"])) / np.max(np.abs(ref["hd"]))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="voronoi64_250nm")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"],
                    help="form whose throughput is the value (mixed: FP32 inner / FP64 outer)")
    ap.add_argument("--multigrid", type=int, default=8,
                    help="most levels of the multigrid preconditioner of the FP32 inner solve (mixed form; "
                         "0: Jacobi)")
    ap.add_argument("--ws-policy", default="measured", choices=["every", "measured"],
                    help="when the network warm-starts a potential solve: every solve, or per solve the "
                         "faster form by measured solve time")
    ap.add_argument("--learned", default="",
                    help="trained warm-start network (.npz); default: the package's network for --config "
                         "when present; 'none' disables")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the plain FP64 run (speedup and errors of the form against it)")
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--general", action="store_true", help="unstructured kernels (SELL K, element loops)")
    ap.add_argument("--inner-tol", type=float, default=0.0)
    ap.add_argument("--amg", action="store_true",
                    help="--general --precision mixed: smoothed-aggregation AMG preconditioner of the FP32 inner solve")
    ap.add_argument("--no-secant", action="store_true",
                    help="chosen form without the secant start of the line search's later potential solves")
    ap.add_argument("--eval-precision", default="fp64", choices=["fp64", "mixed"],
                    help="mixed form: arithmetic of the energy evaluation's C m, D m and G u "
                         "(the inner PCG and truncated CG are FP32 either way)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="after the run, copy 512 MiB on the device (calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-full-step", action="store_true",
                    help="only the CPU baseline, plus one complete oracle field step against its unit-cost model")
    ap.add_argument("--loop", action="store_true",
                    help="instead of the timed window: the chosen form and plain FP64 through the reversal")
    ap.add_argument("--loop-after", type=int, default=4, help="--loop: steps run after the reversal")
    ap.add_argument("--loop-window", type=int, default=24, help="--loop: saturated steps timed separately")
    ap.add_argument("--loop-out", default="", help="--loop: also write the per-step records (JSON) here")
    ap.add_argument("--loop-perturb", type=float, default=0.0,
                    help="--loop: also run plain FP64 from the initial m perturbed by this relative size")
    ap.add_argument("--loop-fp64-only", action="store_true", help="--loop: plain FP64 runs only")
    ap.add_argument("--avalanche", type=int, default=0,
                    help="step index K: FP64 to step K - 1, then step K from that state by FP64 (exact and "
                         "perturbed by 1e-15) and by the chosen form")
    ap.add_argument("--avalanche-seeds", type=int, default=2)
    ap.add_argument("--precond-eval", action="store_true",
                    help="the learned preconditioner against Jacobi and multigrid on potential solves (use --seed "
                         "for a microstructure outside the training set)")
    ap.add_argument("--precond", default="", help="--precond-eval: the preconditioner network (.npz)")
    ap.add_argument("--precond-steps", type=int, default=4)
    ap.add_argument("--precond-every", type=int, nargs="+", default=[16, 64, 256, 1000000],
                    help="--precond-eval: iterations between the network's Galerkin steps (a large value: "
                         "at the start only)")
    ap.add_argument("--reversal-check", action="store_true",
                    help="the chosen form sweeps until the first step with M.h < 0; plain FP64 re-solves that step "
                         "and the one before from the chosen form's states")
    ap.add_argument("--rc-hstart", type=float, default=None, help="--reversal-check: first field (T)")
    ap.add_argument("--rc-hstep", type=float, default=None, help="--reversal-check: field step (T)")
    ap.add_argument("--rc-time-limit", type=float, default=700.0, help="--reversal-check: seconds of sweep at most")
    ap.add_argument("--rc-fine-from", type=float, default=None,
                    help="--reversal-check: below this field (T) the sweep steps by --rc-hstep-fine")
    ap.add_argument("--rc-hstep-fine", type=float, default=-0.02, help="--reversal-check: the fine field step (T)")
    ap.add_argument("--rc-stop-step", type=int, default=None, help="--reversal-check: last step of the sweep")
    ap.add_argument("--rc-resolve", type=int, nargs="+", default=None,
                    help="--reversal-check: the steps plain FP64 re-solves from the chosen form's states (instead of "
                         "the reversal step and the one before)")
    ap.add_argument("--rc-fp64", action="store_true", help="--reversal-check: the FP64 re-solves")
    ap.add_argument("--avalanche-eps", default="1e-15",
                    help="comma-separated relative sizes of the perturbations of m (each with every seed)")
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(args.config, args.cpu_worker, args.seed, full_step=args.cpu_full_step)
        return 0
    if args.cpu_full_step:   # the CPU baseline alone, with one complete oracle field step
        fd, path = tempfile.mkstemp(suffix=".json", prefix="mumag_cpu_")
        os.close(fd)
        env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
        rc = subprocess.call([sys.executable, str(ROOT / "bench.py"), "--cpu-worker", path, "--cpu-full-step",
                              "--config", args.config, "--seed", str(args.seed)], env=env)
        with open(path) as f:
            print(json.dumps(dict(cpu_full_step=args.config, **json.load(f))), flush=True)
        os.unlink(path)
        return rc
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2
    cpu = None
    if not (args.no_cpu_baseline or args.loop or args.avalanche or args.precond_eval or args.reversal_check):
        cpu = start_cpu_leg(args.config, args.seed)
    learned = None
    if args.learned != "none":
        cand = pathlib.Path(args.learned) if args.learned else (
            ROOT / "mammos-mumag_amd" / "mammos_mumag_amd" / "data" / f"warmstart_{args.config}.npz")
        if cand.exists() and not args.general:
            learned = str(cand)
        elif args.learned:
            raise SystemExit(f"bench.py: trained warm-start network {cand} not found")
    if args.loop:
        return loop_main(args, learned)
    if args.avalanche:
        return avalanche_main(args, learned)
    if args.reversal_check:
        return reversal_check_main(args)
    if args.precond_eval:
        if not args.precond:
            args.precond = str(ROOT / "mammos-mumag_amd" / "mammos_mumag_amd" / "data" / f"precond_{args.config}.npz")
        return precond_eval_main(args)
    mgl = args.multigrid if args.precision == "mixed" and not args.general else 0
    amg = bool(args.amg and args.general and args.precision == "mixed")
    res = gpu_run(args, args.precision, traffic_probe=args.traffic_probe, learned=learned, multigrid=mgl, amg=amg)
    ablate = None
    if learned and not args.no_compare:
        ablate = gpu_run(args, args.precision, multigrid=mgl)   # the same form without the network
    jacobi = None
    if args.precision == "mixed" and (mgl or amg) and not args.no_compare:
        jacobi = gpu_run(args, "mixed")   # FP32 inner solve alone: the Jacobi preconditioner
    ref = None
    if (args.precision != "fp64" or learned) and not args.no_compare:
        ref = gpu_run(args, "fp64")
    steps = args.steps
    value = steps / res["elapsed"]
    roof = roofline(res, args.precision, args)
    base = ref if ref is not None else res
    st = base["stats"]   # the CPU baseline runs the plain FP64 method: scale by its counts
    cpu_res = None
    if cpu is not None:
        proc, path = cpu
        proc.wait()
        try:
            with open(path) as f:
                c = json.load(f)
            per_step = (st["pcg_iters"] * c["t_pcg_iter"] + st["fcalls"] * c["t_eval"]
                        + st["tcg_iters"] * c["t_tcg_iter"]) / steps
            cpu_res = dict(value=round(1.0 / per_step, 6), unit="field steps/s", cores=c["cores"],
                           kind="port", host_cores=c["host_cores"], threads_env=c["threads_env"],
                           sample=(f"FP64 numpy/scipy oracle on the same {args.config} mesh: unit costs "
                                   f"timed over {c['n_pcg']} PCG iterations, 3 PCG-free evaluations and "
                                   f"{c['n_tcg']} truncated-CG iterations ({c['sample_s']:.1f} s), "
                                   f"scaled by the plain FP64 run's per-step counts; single thread"))
        except Exception as e:  # the baseline is reported, not required
            log(f"[bench] cpu baseline failed: {e}")
        finally:
            os.unlink(path)
    from mammos_mumag_amd.generate import CONFIGS
    cfgd = CONFIGS[args.config]
    out = {
        "metric": METRIC, "value": round(value, 6), "unit": "field steps/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(1e3 * res["elapsed"] / steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {cfgd['grains']}-grain Voronoi cube {cfgd['edge']:g} nm, "
                               f"{cfgd['n']}^3 voxels x 6 Kuhn tets in a graded air box, "
                               f"{res['n_mag_nodes']} magnet nodes / {res['n_nodes']} nodes / "
                               f"{res['n_elems']} tets, field steps from {cfgd['hstart']:+g} T by {cfgd['hstep']:g} T",
                   "precision": args.precision,
                   "form": (("learned warm start + " if learned else "")
                            + ("FP32 inner solve" + (f" (multigrid preconditioner, levels {res['mg_levels']})" if mgl else "")
                               + (f" (smoothed-aggregation AMG preconditioner, levels {res['amg_levels']}, set-up "
                                  f"{res['amg_setup_s']:.2f} s outside the timed steps)" if amg else "")
                               + " + FP64 outer" if args.precision == "mixed" else "FP64")
                            + (" + secant start of the line search's later solves" if (mgl or amg) and not args.no_secant
                               else "")),
                   "learned": (None if not learned else dict(
                       network=os.path.relpath(learned, ROOT), policy=args.ws_policy,
                       solves=res["warm_start"]["solves"], warm_started_solves=res["warm_start"]["calls"],
                       mean_ms_per_solve_plain=round(1e3 * res["warm_start"]["mean_s_plain"], 3),
                       mean_ms_per_solve_warm=round(1e3 * res["warm_start"]["mean_s_warm"], 3),
                       without_network=None if ablate is None else dict(
                           value=round(steps / ablate["elapsed"], 6),
                           pcg_iters_per_step=round(ablate["stats"]["pcg_iters"] / steps, 2),
                           gain=round(ablate["elapsed"] / res["elapsed"], 4)))),
                   "per_step": {k: round(v / steps, 2) for k, v in res["stats"].items()}},
        "mixed_jacobi": (None if jacobi is None else dict(
            value=round(steps / jacobi["elapsed"], 6),
            pcg_iters_per_step=round(jacobi["stats"]["pcg_iters"] / steps, 2),
            speedup_over_fp64=None if ref is None else round(ref["elapsed"] / jacobi["elapsed"], 4),
            errors=None if ref is None else form_errors(jacobi, ref))),
        "vs_fp64": (None if ref is None else dict(
            fp64_value=round(steps / ref["elapsed"], 6), speedup=round(ref["elapsed"] / res["elapsed"], 4),
            fp64_per_step={k: round(v / steps, 2) for k, v in ref["stats"].items()},
            errors=form_errors(res, ref))),
        "roofline": roof,
        "whole_step": dict(chosen=whole_step(res, args.precision, steps),
                           fp64=None if ref is None else whole_step(ref, "fp64", steps),
                           mixed_jacobi=None if jacobi is None else whole_step(jacobi, "mixed", steps)),
        "cpu_baseline": cpu_res,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
