"""``hysteresis.run``: the reference's one-call hysteresis loop (src/mammos_mumag/hysteresis.py:27-141)
on the GPU path.

The reference takes mammos-entity quantities; this mirror takes plain SI floats (Ms in A/m,
A in J/m, K1 in J/m^3, fields in A/m) and converts them the way mammos-units does through
astropy's ``magnetic_flux_field`` equivalency, B = mu0 H with astropy's CODATA-2018
mu0 = 1.25663706212e-6 (not the scripts' 4 pi 1e-7, nor scipy's CODATA value that the
outputs divide by).  It writes the three-domain ``.krn`` and the ``.p2`` that
``hysteresis.run`` writes (hysteresis.py:56-110: magnet, free space, shell; initial
m = +z; field 1 degree off z; mstep 0.4, mfinal -1.2), runs ``Simulation.run_loop`` and
reads ``hystloop.dat`` back (hysteresis.py:111-141).
"""
import pathlib
import shutil
from dataclasses import dataclass, field

import numpy as np

from .simulation import Simulation

MU0_ASTROPY = 1.25663706212e-6


@dataclass(frozen=True)
class Result:
    """H (A/m), M (A/m), energy density (J/m^3), configuration_type (the snapshot counter of
    each record) and the saved configurations {index: path}."""
    H: np.ndarray
    M: np.ndarray
    energy_density: np.ndarray
    configuration_type: np.ndarray
    configurations: dict = field(default_factory=dict)


def _krn(Js, A, K1):
    """templates/krn.jinja for the three domains of hysteresis.py:64-92."""
    return (f"0 0.0 {K1} 0.0 {Js} {A} 1.0 0.0\n"
            "0.0 0.0 0.0 0.0 0.0 0.0 1.0 0.0\n"
            "0.0 0.0 0.0 0.0 0.0 0.0 1.0 0.0\n")


def _p2(hstart, hfinal, hstep):
    """templates/p2.jinja with the parameters of hysteresis.py:94-110."""
    return "\n".join([
        "[mesh]", "size = 1e-09", "scale = 0", "", "[initial state]", "state = mxyz",
        "mx = 0", "my = 0", "mz = 1", "", "[field]", "hmag_on = 1",
        f"hstart = {hstart!r}", f"hfinal = {hfinal!r}", f"hstep = {hstep!r}",
        "hx = 0.01745", "hy = 0", "hz = 0.99984", "mstep = 0.4", "mfinal = -1.2", "",
        "[minimizer]", "iter_max = 1000", "tol_fun = 1e-10", "tol_hmag_factor = 1",
        "tol_u = 1e-10", "precond_iter = 10", ""])


def run(Ms, A, K1, mesh_filepath, hstart, hfinal, hstep=None, hnsteps=20, outdir="hystloop"):
    """Half a hysteresis loop from ``hstart`` to ``hfinal`` (A/m) on ``mesh_filepath`` (.fly)."""
    if hstep is None:
        hstep = (hfinal - hstart) / hnsteps
    outdir = pathlib.Path(outdir)
    outdir.mkdir(parents=True, exist_ok=True)
    stage = outdir / "_inputs"
    stage.mkdir(exist_ok=True)
    (stage / "hystloop.krn").write_text(_krn(Ms * MU0_ASTROPY, A, K1))
    (stage / "hystloop.p2").write_text(_p2(hstart * MU0_ASTROPY, hfinal * MU0_ASTROPY, hstep * MU0_ASTROPY))
    shutil.copyfile(mesh_filepath, stage / "hystloop.fly")
    sim = Simulation(mesh_filepath=stage / "hystloop.fly", materials_filepath=stage / "hystloop.krn",
                     parameters_filepath=stage / "hystloop.p2")
    sim.run_loop(outdir=outdir, name="hystloop")
    dat = np.loadtxt(outdir / "hystloop.dat", ndmin=2)
    vtus = sorted(p.resolve() for p in outdir.glob("*.vtu"))
    return Result(H=dat[:, 1] / MU0_ASTROPY, M=dat[:, 2] / MU0_ASTROPY, energy_density=dat[:, 3],
                  configuration_type=dat[:, 0].astype(np.int64),
                  configurations={i + 1: p for i, p in enumerate(vtus)})
