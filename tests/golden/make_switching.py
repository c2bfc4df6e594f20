"""Write the switching fixtures under tests/golden/.

1. ``quickstart_loop.json``: the reversal loop that the reference records in
   ``examples/quickstart.ipynb`` (cells 4 and 10-14: ``hysteresis.run`` on
   ``mesh.CUBE_20_nm``, which is byte-identical to ``tests/data/cube.fly``).  The
   notebook is read as JSON data; the printed arrays H (A/m), M (A/m), energy density
   (J/m^3) and configuration_type are parsed from the cell outputs.  The inputs that
   ``hysteresis.run`` writes (src/mammos_mumag/hysteresis.py:56-110) are reconstructed
   with the unit conversions of mammos-units (astropy's magnetic_flux_field equivalency,
   B = mu0 H with astropy's CODATA-2018 mu0 = 1.25663706212e-6, which is not installed
   here): Js = Ms mu0, hstart = (10 T -> A/m -> T), hstep = ((hfinal - hstart) / 20 in A/m
   -> T).  They are also written as ``quickstart.krn`` / ``quickstart.p2`` in the layout of
   the reference's templates (templates/krn.jinja, templates/p2.jinja).

2. ``oracle_branches.json``: the FP64 oracle's per-step records and counters on the
   cases that take the minimiser through its rarely used branches, for the GPU tests:

   * ``quickstart``: the loop above (reversal at record 16; no backtracks, no restarts);
   * ``tilt1``: the same magnet with the easy axis tilted by 1 rad from z (theta = 1),
     +10 T to -10 T in -2 T steps: 33 descent restarts (16 to -g, 17 to -z) at 0 T, and
     reversal at -4 T;
   * ``vortex_hmag0`` / ``vortex_hmag1``: one field step at 0 T from the vortex initial
     state (scripts/magnetization.py:24-33) without / with magnetostatics: 8 and 1
     line-search halvings.

   Every restart and halving decision of these cases is far from its threshold (the
   smallest margin of the restart test is 0.011 in cosine, of a halving test 0.07 of |f|),
   so an FP64 implementation with another summation order takes the same branches.  The
   stop of a slowly converging NCG is not: its tests hover near their thresholds, and
   rounding moves it by a few iterations.  Each case therefore also records ``spread``:
   the oracle rerun from initial states perturbed by 1e-15 relative (3 seeds), with the
   range of its final counters and the largest change of M.h and of E / |E| over the
   records — the differences that rounding alone produces in the oracle itself.

Run: ``python tests/golden/make_switching.py`` (the notebook part needs
/root/reference present; ``--oracle-only`` rewrites only the oracle records).
"""
import json
import math
import pathlib
import re
import sys

import numpy as np

DST = pathlib.Path(__file__).resolve().parent
ROOT = DST.parent.parent
NOTEBOOK = pathlib.Path("/root/reference/examples/quickstart.ipynb")
MU0_ASTROPY = 1.25663706212e-6   # CODATA 2018 (astropy.constants.mu0)

QS = dict(Ms=1280000.0, A=7.7e-12, K1=4300000.0, hstart_T=10.0, hfinal_T=-10.0, hnsteps=20,
          h=[0.01745, 0.0, 0.99984], mstep=0.4, mfinal=-1.2, tol_fun=1e-10, tol_hmag_factor=1.0,
          precond_iter=10)


def quickstart_inputs():
    """The .krn / .p2 values hysteresis.run writes for the notebook's call."""
    hstart_Am = QS["hstart_T"] / MU0_ASTROPY
    hfinal_Am = QS["hfinal_T"] / MU0_ASTROPY
    hstep_Am = (hfinal_Am - hstart_Am) / QS["hnsteps"]
    return dict(Js=QS["Ms"] * MU0_ASTROPY, A=QS["A"], K1=QS["K1"], hstart=hstart_Am * MU0_ASTROPY,
                hfinal=hfinal_Am * MU0_ASTROPY, hstep=hstep_Am * MU0_ASTROPY, h=QS["h"],
                mstep=QS["mstep"], mfinal=QS["mfinal"], tol_fun=QS["tol_fun"],
                tol_hmag_factor=QS["tol_hmag_factor"], precond_iter=QS["precond_iter"])


def krn_text(Js, A, K1, theta=0, phi=0.0):
    """templates/krn.jinja: theta phi K1 K2 Js A 1.0 0.0, one line per domain."""
    lines = [f"{theta} {phi} {K1} 0.0 {Js} {A} 1.0 0.0"]
    lines += ["0.0 0.0 0.0 0.0 0.0 0.0 1.0 0.0"] * 2
    return "\n".join(lines) + "\n"


def p2_text(q, m=(0, 0, 1)):
    """templates/p2.jinja with the values hysteresis.run passes (hysteresis.py:94-110)."""
    return "\n".join([
        "[mesh]", "size = 1e-09", "scale = 0", "", "[initial state]", "state = mxyz",
        f"mx = {m[0]}", f"my = {m[1]}", f"mz = {m[2]}", "", "[field]", "hmag_on = 1",
        f"hstart = {q['hstart']!r}", f"hfinal = {q['hfinal']!r}", f"hstep = {q['hstep']!r}",
        f"hx = {q['h'][0]}", f"hy = {q['h'][1]}", f"hz = {q['h'][2]}", f"mstep = {q['mstep']}",
        f"mfinal = {q['mfinal']}", "", "[minimizer]", "iter_max = 1000", f"tol_fun = {q['tol_fun']}",
        f"tol_hmag_factor = {q['tol_hmag_factor']}", f"tol_u = {q['tol_fun'] * q['tol_hmag_factor']}",
        f"precond_iter = {q['precond_iter']}", ""])


def _array_after(text):
    body = text[text.index("[") + 1:text.index("]")]
    return [float(v) for v in body.replace("\n", " ").split(",") if v.strip()]


def parse_notebook(path=NOTEBOOK):
    nb = json.loads(path.read_text())
    out = {}
    for cell in nb["cells"]:
        src = "".join(cell.get("source", [])).strip()
        key = {"hysteresis_result.H": "H", "hysteresis_result.M": "M",
               "hysteresis_result.energy_density": "energy_density",
               "hysteresis_result.configuration_type": "configuration_type"}.get(src)
        if key is None:
            continue
        text = "".join(cell["outputs"][0]["data"]["text/plain"])
        out[key] = _array_after(text)
        if key != "configuration_type":
            out[key + "_unit"] = re.search(r"unit='([^']*)'", text).group(1)
    out["configuration_type"] = [int(v) for v in out["configuration_type"]]
    return out


# ---------------------------------------------------------------- oracle cases

def _norm(v):
    s = math.sqrt(sum(x * x for x in v))
    return [x / s for x in v]


def vortex_state(model):
    """scripts/magnetization.py:24-33 (state 'vortex'), masked to the magnet, normalised."""
    x = model.coords
    r = np.sqrt(x[:, 1] ** 2 + x[:, 2] ** 2)
    rc = 0.14 * r.max()
    m = np.zeros((model.N, 3))
    with np.errstate(divide="ignore", invalid="ignore"):
        m[:, 0] = np.exp(-2 * (r / rc))
        m[:, 1] = -(x[:, 2] / r) * np.sqrt(1 - np.exp(-4 * (r * r) / (rc * rc)))
        m[:, 2] = (x[:, 1] / r) * np.sqrt(1 - np.exp(-4 * (r * r) / (rc * rc)))
    m = np.nan_to_num(m)
    m[model.meas <= 0] = 0.0
    ln = np.linalg.norm(m, axis=1)
    z = (ln == 0).astype(np.float64)
    return (m * ((1.0 - z) / (ln + z))[:, None]).ravel()


def branch_cases():
    """name -> (krn dict of tag -> material, params dict, initial state name, max_steps)."""
    q = quickstart_inputs()
    air = dict(theta=0.0, phi=0.0, K1=0.0, Js=0.0, A=0.0)

    def krn(theta=0.0):
        return {1: dict(theta=theta, phi=0.0, K1=q["K1"], Js=q["Js"], A=q["A"]), 2: air, 3: air}

    base = dict(m=[0.0, 0.0, 1.0], mstep=q["mstep"], mfinal=q["mfinal"], state_id=0, hmag_on=1,
                tol_u=q["tol_fun"] * q["tol_hmag_factor"], tol_fun=q["tol_fun"],
                precond_iter=q["precond_iter"], iter_max=1000, h=_norm(q["h"]))
    return {
        "quickstart": (krn(), dict(base, hstart=q["hstart"], hfinal=q["hfinal"], hstep=q["hstep"]),
                       "uniform", None),
        "tilt1": (krn(1.0), dict(base, hstart=10.0, hfinal=-10.0, hstep=-2.0), "uniform", None),
        "vortex_hmag0": (krn(), dict(base, hmag_on=0, hstart=0.0, hfinal=-1.0, hstep=-1.0), "vortex", 1),
        "vortex_hmag1": (krn(), dict(base, hstart=0.0, hfinal=-1.0, hstep=-1.0), "vortex", 1),
    }


COUNTERS = ("ncg_iters", "calls", "pcg_iters", "tcg_iters", "backtracks", "restarts_g", "restarts_z")


def run_oracle_case(name, coords, tets, etags, perturb_seed=None):
    """The oracle's loop for case ``name``; ``perturb_seed`` scales the initial state by
    (1 + 1e-15 N(0, 1)) per entry first."""
    sys.path.insert(0, str(ROOT))
    from oracle.assemble import Model
    from oracle.loop import run_loop
    krn, params, state, max_steps = branch_cases()[name]
    model = Model(coords, tets, etags, krn)
    m0 = model.uniform_m(params["m"]) if state == "uniform" else vortex_state(model)
    if perturb_seed is not None:
        m0 = m0 * (1 + 1e-15 * np.random.default_rng(perturb_seed).normal(size=m0.shape))
    steps = []
    rec, mini = run_loop(model, params, m0, max_steps=max_steps,
                         on_step=lambda i, h, mh, E, st: steps.append({k: int(st.get(k, 0)) for k in COUNTERS}))
    return dict(records=[[float(v) for v in r] for r in rec], counters=steps)


def spread(name, base, coords, tets, etags, seeds=(1, 2, 3)):
    runs = [base] + [run_oracle_case(name, coords, tets, etags, s) for s in seeds]
    fin = [r["counters"][-1] for r in runs]
    rb = np.array(base["records"])
    out = {k: [min(f[k] for f in fin), max(f[k] for f in fin)] for k in COUNTERS}
    out["n_records"] = sorted({len(r["records"]) for r in runs})
    same = [np.array(r["records"]) for r in runs if len(r["records"]) == len(rb)]
    out["mh"] = float(max(np.max(np.abs(r[:, 2] - rb[:, 2])) for r in same))
    out["E_rel"] = float(max(np.max(np.abs(r[:, 3] / rb[:, 3] - 1)) for r in same))
    return out


def main(argv):
    z = np.load(DST / "cube_mesh.npz")
    coords, tets, etags = z["coords"], z["tets"], z["elem_tags"]
    if "--oracle-only" not in argv:
        q = quickstart_inputs()
        nb = parse_notebook()
        nb["inputs"] = q
        nb["source"] = "examples/quickstart.ipynb cells 4, 10-14 (hysteresis.run on mesh.CUBE_20_nm)"
        nb["mu0_astropy"] = MU0_ASTROPY
        (DST / "quickstart_loop.json").write_text(json.dumps(nb, indent=1) + "\n")
        (DST / "quickstart.krn").write_text(krn_text(q["Js"], q["A"], q["K1"]))
        (DST / "quickstart.p2").write_text(p2_text(q))
    out = {}
    for name in branch_cases():
        out[name] = run_oracle_case(name, coords, tets, etags)
        out[name]["spread"] = spread(name, out[name], coords, tets, etags)
    (DST / "oracle_branches.json").write_text(json.dumps(out, indent=1) + "\n")
    for name, r in out.items():
        print(name, len(r["records"]), r["counters"][-1], r["spread"])


if __name__ == "__main__":
    main(sys.argv[1:])
