"""Write ``sphere_shell.json``: FP64 oracle results on generated spherical-shell meshes
(``generate.build_sphere``), which take the reference's shell branch (scripts/matrix.py:49-70
with the Imhoff tensor of scripts/trafo.py:24-45).

* ``loop``: the demagnetisation loop of a 4-grain 20 nm cube (16 voxels per edge, 5 free
  space and 4 shell layers per side: 42,875 nodes, 235,824 tets) from 0 T down in -1 T
  steps, with its per-step records [counter, h, M.h, E] and counters, and ``spread``: the
  oracle rerun from initial states perturbed by 1e-15 relative (2 seeds) — the largest
  change of M.h and of E / |E| that rounding alone produces.
* ``hmag_uniform``: the magnetostatic energy of the uniformly magnetised single-grain cube
  (m = +z, u0 = 0, PCG tolerance 1e-10) on meshes of 16, 24 and 60 voxels per edge (up to
  704,969 nodes and 4,088,832 tets), with the PCG iteration count.  ``analytic`` is the
  energy density of a uniformly magnetised cube, Js^2 / 6 in the units of
  scripts/energies.py (energies scaled by mu0; demagnetising factor 1/3).

Run: ``python tests/golden/make_sphere.py`` (about 5 minutes on one core).
"""
import json
import pathlib
import sys
import time

import numpy as np

DST = pathlib.Path(__file__).resolve().parent
ROOT = DST.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "mammos-mumag_amd"))

LOOP_MESH = dict(edge=20.0, n=16, air_layers=5, shell_layers=4, grains=4, cone=15.0, seed=5)
LOOP_PARAMS = dict(m=[0.0, 0.0, 1.0], state_id=0, mstep=0.4, mfinal=-1.2, hstart=0.0, hfinal=-8.0,
                   hstep=-1.0, hmag_on=1, tol_fun=1e-10, tol_u=1e-10, precond_iter=10, iter_max=1000,
                   verbose=0)
UNIFORM_MESHES = {"16": dict(edge=20.0, n=16, air_layers=5, shell_layers=4),
                  "24": dict(edge=20.0, n=24, air_layers=6, shell_layers=5),
                  "60": dict(edge=20.0, n=60, air_layers=8, shell_layers=6)}
COUNTERS = ("ncg_iters", "calls", "pcg_iters", "tcg_iters", "backtracks", "restarts_g", "restarts_z")


def loop_params():
    from mammos_mumag_amd.fileio import _normalize
    from mammos_mumag_amd.generate import FIELD_DIR
    return dict(LOOP_PARAMS, h=_normalize(list(FIELD_DIR)))


def run_loop_case(perturb_seed=None):
    from mammos_mumag_amd import generate
    from oracle.assemble import Model
    from oracle.loop import run_loop
    mesh = generate.build_sphere(**LOOP_MESH)
    model = Model(mesh.coords(), mesh.tets(), mesh.elem_tags(), mesh.krn())
    params = loop_params()
    m0 = model.uniform_m(params["m"])
    if perturb_seed is not None:
        m0 = m0 * (1 + 1e-15 * np.random.default_rng(perturb_seed).normal(size=m0.shape))
    steps = []
    rec, _ = run_loop(model, params, m0,
                      on_step=lambda i, h, mh, E, st: steps.append({k: int(st.get(k, 0)) for k in COUNTERS}))
    return dict(records=[[float(v) for v in r] for r in rec], counters=steps)


def main():
    from mammos_mumag_amd import generate
    from oracle.assemble import Model
    from oracle.energies import hmag_eg
    t0 = time.time()
    out = dict(loop_mesh=LOOP_MESH, loop_params=loop_params(), analytic=1.61 ** 2 / 6.0)
    base = run_loop_case()
    runs = [run_loop_case(s) for s in (1, 2)]
    rb = np.array(base["records"])
    same = [np.array(r["records"]) for r in runs if len(r["records"]) == len(rb)]
    fin = [r["counters"][-1] for r in [base] + runs]
    base["spread"] = dict(
        n_records=sorted({len(r["records"]) for r in [base] + runs}),
        mh=float(max(np.max(np.abs(r[:, 2] - rb[:, 2])) for r in same)),
        E_rel=float(max(np.max(np.abs(r[:, 3] / rb[:, 3] - 1)) for r in same)),
        **{k: [min(f[k] for f in fin), max(f[k] for f in fin)] for k in COUNTERS})
    out["loop"] = base
    print("loop", len(base["records"]), base["counters"][-1], base["spread"], f"{time.time() - t0:.0f} s", flush=True)
    out["hmag_uniform"] = {}
    for key, kw in UNIFORM_MESHES.items():
        mesh = generate.build_sphere(**kw)
        model = Model(mesh.coords(), mesh.tets(), mesh.elem_tags(), mesh.krn())
        st = {}
        E, _, _ = hmag_eg(model, model.uniform_m([0.0, 0.0, 1.0]), np.zeros(model.N), 1e-10, st)
        out["hmag_uniform"][key] = dict(mesh=kw, N=int(model.N), NE=int(len(model.tets)), E=float(E),
                                        pcg_iters=int(st["pcg_iters"]), R=float(model.R), Rinf=float(model.Rinf))
        print(key, out["hmag_uniform"][key], f"{time.time() - t0:.0f} s", flush=True)
        del model, mesh
    (DST / "sphere_shell.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
