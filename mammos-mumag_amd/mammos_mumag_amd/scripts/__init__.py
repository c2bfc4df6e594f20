"""Drop-in entry points: ``python -m mammos_mumag_amd.scripts.<script> <name>``.

Each runs in the working directory that holds ``<name>.fly/.krn[/.p2]`` and writes the
same output files as the reference's escript scripts (``src/mammos_mumag/scripts``):

* ``loop``     -> ``<name>.dat``, ``<name>_NNNN.vtu``, ``<name>_stats.txt`` (loop.py:144-174)
* ``hmag``     -> ``<name>.csv``, ``<name>_hmag.vtu`` (hmag.py:111-165)
* ``exani``    -> ``<name>_vortex.csv``, ``<name>_uniform.csv`` (exani.py:36-87)
* ``external`` -> ``<name>.csv`` (external.py:44-74)
* ``mapping``  -> ``<name>_{zeeman,exchange,anisotropy,hmag,energy}.csv``, ``<name>_stats.txt``
  (mapping.py:22-118, 238-295)
"""
import gc
import inspect
import math
import os
import sys
import time

import numpy as np

from ..fileio import mu0_output, read_p2, write_csv, write_vtu


def _rss_mb():
    try:
        import psutil
        return psutil.Process(os.getpid()).memory_info().rss / (1024 * 1024)
    except Exception:
        return float("nan")


def _name(argv, script):
    try:
        return argv[1]
    except IndexError:
        sys.exit(f"usage: python -m mammos_mumag_amd.scripts.{script} modelname")


def _model(name):
    """The device context of ``<name>``; ``MAMMOS_MUMAG_PRECISION`` = ``fp64`` (default) or
    ``mixed`` selects the arithmetic.  The mixed form preconditions its FP32 potential solve
    with the smoothed-aggregation AMG when the mesh has more nodes than the AMG's coarsest
    level (500; ``MAMMOS_MUMAG_AMG=0`` keeps Jacobi) and starts the line search's later
    potential solves from the secant step.  These change iteration counts, not the solve's
    tolerance; ``fp64`` keeps the reference's plain method and counts."""
    from ..model import Model
    precision = os.environ.get("MAMMOS_MUMAG_PRECISION", "fp64")
    model = Model.from_files(name, precision=precision)
    if precision == "mixed":
        if os.environ.get("MAMMOS_MUMAG_AMG", "1") != "0" and model.N > 500:
            model.amg()
        model.secant_start(True)
    return model


def _krn_first(name):
    """First .krn line (scripts/tools.py:39-62 read_Js / read_A / readAnisotropyEnergy)."""
    with open(name + ".krn") as f:
        ll = f.readline().split()
    return dict(theta=float(ll[0]), phi=float(ll[1]), K1=float(ll[2]), Js=float(ll[4]), A=float(ll[5]))


def _aniso_energy(name, m):
    """scripts/tools.py:51-62 (uses the script mu0 = 4 pi 1e-7)."""
    n = math.sqrt(sum(v * v for v in m))
    m = [v / n for v in m] if n else m
    k = _krn_first(name)
    n0 = math.sin(k["theta"]) * math.cos(k["phi"])
    n1 = math.sin(k["theta"]) * math.sin(k["phi"])
    n2 = math.cos(k["theta"])
    return -4.0e-7 * math.pi * k["K1"] * (m[0] * n0 + m[1] * n1 + m[2] * n2) ** 2.0


def main_loop(argv):
    from ..loop import run_loop
    name = _name(argv, "loop")
    params = read_p2(name + ".p2")
    pre = _rss_mb()
    model = _model(name)
    post = _rss_mb()
    gc.collect()
    col = _rss_mb()
    with open(name + "_stats.txt", "w") as f:
        f.write(inspect.cleandoc(f"""
            Memory before escript2jax: {pre} MB.
            Memory after  escript2jax: {post} MB.
            Memory after garbage collection: {col} MB.
            """))
    m0 = model.state_m(params["m"], params["state_id"])
    run_loop(model, params, m0, name=name)


def main_exani(argv):
    name = _name(argv, "exani")
    model = _model(name)
    mu0 = mu0_output()
    k0 = _krn_first(name)
    m, k = model.vortex_xm()
    E, _ = model.exani_eg(model.to_device(m))
    write_csv(name + "_vortex.csv", "Exchange energy density of a vortex on a, b plane.", [
        ("E_gradient", E / mu0, "Energy evaluated from gradient (J/m^3)."),
        ("E_analytic", (2 * k * k * mu0 * k0["A"] / (1e-9 * 1e-9)) / mu0,
         "Energy evaluated analytically (J/m^3)."),
    ])
    uniform = [0.0, 0.0, 1.0]
    E, _ = model.exani_eg(model.to_device(model.uniform_m(uniform)))
    write_csv(name + "_uniform.csv",
              f"Magnetocrystalline anisotropy energy density of uniformly magnetized sample in direction {uniform}.",
              [("E_gradient", E / mu0, "Energy evaluated from gradient (J/m^3)."),
               ("E_analytic", _aniso_energy(name, uniform) / mu0, "Energy evaluated analytically (J/m^3).")])


def main_external(argv):
    name = _name(argv, "external")
    p = read_p2(name + ".p2")
    model = _model(name)
    mu0 = mu0_output()
    value = p["hstart"] - p["hstep"]
    m, h = p["m"], p["h"]
    Js = _krn_first(name)["Js"]
    ezee = -Js * value * (m[0] * h[0] + m[1] * h[1] + m[2] * h[2])
    E = model.external_eg(model.to_device(model.uniform_m(m)), value, h)
    write_csv(name + ".csv", f"Zeeman energy density for an external field of mu_0 Hext {value} (T).", [
        ("E_gradient", E / mu0, "Energy evaluated from gradient (J/m^3)."),
        ("E_analytic", ezee / mu0, "Energy evaluated analytically (J/m^3)."),
    ])


def main_hmag(argv):
    name = _name(argv, "hmag")
    model = _model(name)
    mu0 = mu0_output()
    Js = _krn_first(name)["Js"]
    m = model.to_device(model.uniform_m([0.0, 0.0, 1.0]))
    u = model.empty1()
    E, g, _ = model.hmag_eg(m, u, 1e-12)
    b = model.divergence(m)
    E_field = 0.5 * model.dot(b, u) / model.V  # = -1/2 int h . Js m / V (hmag.py:131)
    write_csv(name + ".csv", "Magnetostatic energy density of uniformly magnetized cube.", [
        ("E_field", E_field / mu0, "Energy density evaluated from field (J/m^3)."),
        ("E_gradient", E / mu0, "Energy density evaluated from gradient (J/m^3)."),
        ("E_analytic", (Js * Js / 6) / mu0, "Energy density evaluated analytically (J/m^3)."),
    ])
    uh = u.cpu().numpy()
    gh = g.cpu().numpy().reshape(-1, 3)
    meas = np.where(model.meas == 0, 1.0, model.meas)
    h_nodes = -model.V * gh / meas[:, None]
    # cell field h = -grad u (hmag.py:40-42)
    X = model.coords[model.tets]
    J = X[:, 1:, :] - X[:, 0:1, :]
    du = uh[model.tets[:, 1:]] - uh[model.tets[:, 0:1]]
    h_cell = -np.linalg.solve(J, du[..., None])[..., 0]
    write_vtu(name + "_hmag.vtu", model.coords, model.tets, model.elem_tags,
              {"m": m.cpu().numpy().reshape(-1, 3), "U": uh, "h_nodes": h_nodes},
              cell_data={"h": h_cell})


def main_mapping(argv):
    name = _name(argv, "mapping")
    p = read_p2(name + ".p2")
    pre = _rss_mb()
    model = _model(name)
    post = _rss_mb()
    gc.collect()
    col = _rss_mb()
    mu0 = mu0_output()
    k0 = _krn_first(name)
    m_host = model.state_m(p["m"], p["state_id"])
    m = model.to_device(m_host)
    value = p["hstart"] - p["hstep"]
    mm, h = p["m"], p["h"]
    total = 0.0
    # Zeeman (mapping.py:22-43)
    E_z = model.external_eg(m, value, h)
    ezee = -k0["Js"] * value * (mm[0] * h[0] + mm[1] * h[1] + mm[2] * h[2])
    write_csv(name + "_zeeman.csv", f"Zeeman energy density for an external field of mu_0 Hext {value} (T).",
              [("E_jax", E_z / mu0, "Energy evaluated with jax backend (J/m^3)."),
               ("E_analytic", ezee / mu0, "Energy evaluated analytically (J/m^3).")])
    total += E_z
    # exchange of the vortex (mapping.py:46-70)
    xm, k = model.vortex_xm()
    E_x, _ = model.exani_eg(model.to_device(xm))
    write_csv(name + "_exchange.csv", "Exchange energy density of a vortex on a, b plane.",
              [("E_jax", E_x / mu0, "Energy evaluated with jax backend (J/m^3)."),
               ("E_analytic", (2 * k * k * mu0 * k0["A"] / (1e-9 * 1e-9)) / mu0,
                "Energy evaluated analytically (J/m^3).")])
    # anisotropy of the initial state (mapping.py:72-94)
    uniform = [0.0, 0.0, 1.0]
    E_a, _ = model.exani_eg(m)
    write_csv(name + "_anisotropy.csv",
              f"Magnetocrystalline anisotropy energy density of uniformly magnetized sample in direction {uniform}.",
              [("E_jax", E_a / mu0, "Energy evaluated with jax backend (J/m^3)."),
               ("E_analytic", _aniso_energy(name, uniform) / mu0, "Energy evaluated analytically (J/m^3).")])
    total += E_a
    # magnetostatics (mapping.py:101-118)
    E_h, _, _ = model.hmag_eg(m, model.empty1(), 1e-10)
    write_csv(name + "_hmag.csv", "Magnetostatic energy density of uniformly magnetized cube.",
              [("E_jax", E_h / mu0, "Energy evaluated with jax backend (J/m^3)."),
               ("E_analytic", (k0["Js"] * k0["Js"] / 6) / mu0, "Energy evaluated analytically (J/m^3).")])
    total += E_h
    # total energy at hstart - hstep (mapping.py:252-295)
    t0 = time.time()
    E, _, _, _ = model.total_eg(m, value, h, p["tol_u"], p["hmag_on"], model.empty1())
    elapsed = time.time() - t0
    with open(name + "_stats.txt", "w") as f:
        f.write(inspect.cleandoc(f"""
            MAP FINITE ELEMENT BACKEND (esys-escript) TO JAX.
            Memory before escript2jax: {pre} MB.
            Memory after  escript2jax: {post} MB.
            Memory after garbage collection: {col} MB.
            Timing and statistics.
            elapsed time: {elapsed}
            function_calls: 1
            """))
    write_csv(name + "_energy.csv", "Total energy density of uniformly magnetized state.",
              [("E_jax", E / mu0, "Energy evaluated with jax backend (J/m^3)."),
               ("E_analytic", total / mu0, "Energy evaluated analytically (J/m^3).")])
