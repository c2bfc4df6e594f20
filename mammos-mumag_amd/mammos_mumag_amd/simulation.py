"""``Simulation``: the reference's file-contract API (src/mammos_mumag/simulation.py).

``run_loop / run_hmag / run_exani / run_external / run_mapping (outdir, name)`` copy
``<name>.fly``, ``<name>.krn`` (and ``<name>.p2``) into ``outdir`` and run the
matching entry point in a child process with ``cwd=outdir``
(simulation.py:87-113, 408-431).  The launcher is this Python interpreter running
``mammos_mumag_amd.scripts.<script>`` instead of ``run-escript``.  A non-zero exit
raises ``RuntimeError`` with the child's stderr; ``info.json`` records the run.
"""
import datetime
import json
import os
import pathlib
import shutil
import subprocess
import sys

from . import __version__

_PKG_ROOT = pathlib.Path(__file__).resolve().parent.parent


def check_dir(outdir):
    outdir = pathlib.Path(outdir)
    outdir.mkdir(exist_ok=True, parents=True)
    return outdir


class Simulation:
    def __init__(self, mesh_filepath=None, materials_filepath=None, parameters_filepath=None):
        self.mesh_filepath = None if mesh_filepath is None else pathlib.Path(mesh_filepath)
        self.materials_filepath = None if materials_filepath is None else pathlib.Path(materials_filepath)
        self.parameters_filepath = None if parameters_filepath is None else pathlib.Path(parameters_filepath)

    def check_attribute(self, *args):
        for attr in args:
            if getattr(self, attr) is None:
                raise AttributeError(f"Attribute `{attr}` has not been defined yet.")

    def _stage(self, outdir, name, need_p2):
        attrs = ["mesh_filepath", "materials_filepath"] + (["parameters_filepath"] if need_p2 else [])
        self.check_attribute(*attrs)
        outdir = check_dir(outdir)
        shutil.copyfile(self.mesh_filepath, outdir / f"{name}.fly")
        shutil.copyfile(self.materials_filepath, outdir / f"{name}.krn")
        if need_p2:
            shutil.copyfile(self.parameters_filepath, outdir / f"{name}.p2")
        return outdir

    @classmethod
    def run_script(cls, script, outdir, name):
        env = dict(os.environ)
        env["PYTHONPATH"] = str(_PKG_ROOT) + os.pathsep + env.get("PYTHONPATH", "")
        cmd = [sys.executable, "-m", f"mammos_mumag_amd.scripts.{script}", name]
        res = subprocess.run(cmd, cwd=outdir, stderr=subprocess.PIPE, env=env)
        if res.returncode:
            raise RuntimeError(f"Simulation has failed. Exit with error: \n{res.stderr.decode('utf-8')}")
        with open(pathlib.Path(outdir) / "info.json", "w") as f:
            json.dump({"datetime": datetime.datetime.now(datetime.timezone.utc).astimezone()
                       .isoformat(timespec="seconds"), "mammos_mumag_amd_version": __version__}, f)

    def run_exani(self, outdir="exani", name="out"):
        self.run_script("exani", self._stage(outdir, name, False), name)

    def run_external(self, outdir="external", name="out"):
        self.run_script("external", self._stage(outdir, name, True), name)

    def run_hmag(self, outdir="hmag", name="out"):
        self.run_script("hmag", self._stage(outdir, name, False), name)

    def run_loop(self, outdir="loop", name="out"):
        self.run_script("loop", self._stage(outdir, name, True), name)

    def run_mapping(self, outdir="magnetization", name="out"):
        self.run_script("mapping", self._stage(outdir, name, True), name)
