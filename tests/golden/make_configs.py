"""Write ``config0_loop.json``: the FP64 oracle's hysteresis loop on the single-grain
workload ``cube40_single`` (40 nm cube, 16^3 voxels x 6 Kuhn tets, graded air box of 6
layers per side: 24,389 nodes; Js 1.61 T, A 7.7 pJ/m, K1 4.3 MJ/m^3, easy axis z; m0 = +z;
field 1 degree off z from +8 T to -8 T in -0.1 T steps, 161 steps; tol_fun 1e-10, PCG
1e-10, precond_iter 10; mfinal -2 so the sweep runs past the reversal).

The loop takes a few minutes on one CPU, too long for a test, so its records
(counter, hext, M.h, E) and cumulative counters per step are stored for the GPU test.
``spread``: the same loop from an initial state perturbed by 1e-15 relative, with the
largest change of M.h and of E / |E| over the records and the range of the final
counters — what rounding alone moves in the oracle.

Run: ``python tests/golden/make_configs.py`` (no reference files needed).
"""
import json
import pathlib
import sys

import numpy as np

DST = pathlib.Path(__file__).resolve().parent
ROOT = DST.parent.parent
for _p in (ROOT, ROOT / "mammos-mumag_amd"):
    sys.path.insert(0, str(_p))

COUNTERS = ("ncg_iters", "calls", "pcg_iters", "tcg_iters", "backtracks", "restarts_g", "restarts_z")


def run(perturb_seed=None):
    from mammos_mumag_amd import generate
    from oracle.assemble import Model
    from oracle.loop import run_loop
    mesh = generate.build_config("cube40_single")
    model = Model(mesh.coords(), mesh.tets(), mesh.elem_tags(), mesh.krn())
    params = generate.loop_params("cube40_single")
    m0 = model.uniform_m(params["m"])
    if perturb_seed is not None:
        m0 = m0 * (1 + 1e-15 * np.random.default_rng(perturb_seed).normal(size=m0.shape))
    steps = []
    rec, _ = run_loop(model, params, m0,
                      on_step=lambda i, h, mh, E, st: steps.append({k: int(st.get(k, 0)) for k in COUNTERS}))
    return dict(records=[[float(v) for v in r] for r in rec], counters=steps)


def main():
    base = run()
    other = run(perturb_seed=1)
    rb, ro = np.array(base["records"]), np.array(other["records"])
    n = min(len(rb), len(ro))
    sp = {k: sorted([base["counters"][-1][k], other["counters"][-1][k]]) for k in COUNTERS}
    sp["n_records"] = sorted({len(rb), len(ro)})
    sp["mh"] = float(np.max(np.abs(rb[:n, 2] - ro[:n, 2])))
    sp["E_rel"] = float(np.max(np.abs(rb[:n, 3] / ro[:n, 3] - 1)))
    base["spread"] = sp
    base["workload"] = "cube40_single"
    (DST / "config0_loop.json").write_text(json.dumps(base, indent=1) + "\n")
    sw = int(np.argmax(rb[:, 2] < 0))
    print(len(rb), "records; reversal at step", sw, "hext", rb[sw, 1], base["counters"][-1], sp)


if __name__ == "__main__":
    main()
