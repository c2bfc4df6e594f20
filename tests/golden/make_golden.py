"""Write the golden fixtures under tests/golden/ from the reference's own test data.

Reads (as data files, never as code) ``/root/reference/tests/data``:

* ``cube.fly`` (the one test mesh, 9,698 nodes / 53,275 Tet4) -> ``cube_mesh.npz``
  (coords, 0-based tets, element tags).
* ``cube.krn`` / ``loop/cube.p2`` -> copied verbatim (text inputs).
* ``loop/cube.dat``, ``hmag/cube.csv``, ``exani/cube_{vortex,uniform}.csv``,
  ``external/cube.csv``, ``mapping/cube_*.csv`` -> copied verbatim (expected outputs).
* ``loop/cube_0001.vtu`` -> ``cube_0001_vtu.npz``: point data ``m``, ``u`` and the
  cell order (index of each VTU cell in the .fly element list).

Run: ``python tests/golden/make_golden.py`` (needs /root/reference present).
"""
import pathlib
import shutil

import numpy as np

SRC = pathlib.Path("/root/reference/tests/data")
DST = pathlib.Path(__file__).resolve().parent


def _read_fly(path):
    lines = path.read_text().splitlines()
    n = int(lines[1].split()[1])
    nodes = np.array(" ".join(lines[2:2 + n]).split(), dtype=np.float64).reshape(n, 6)
    ne = int(lines[2 + n].split()[1])
    el = np.array(" ".join(lines[3 + n:3 + n + ne]).split(), dtype=np.int64).reshape(ne, 6)
    assert np.array_equal(nodes[:, 0], np.arange(1, n + 1))
    return nodes[:, 3:6], (el[:, 2:6] - 1).astype(np.int32), el[:, 1].astype(np.int32)


def _vtu_array(lines, name):
    for i, line in enumerate(lines):
        if line.startswith("<DataArray") and f'Name="{name}"' in line:
            j = i + 1
            while not lines[j].startswith("</DataArray>"):
                j += 1
            return " ".join(lines[i + 1:j]).split()
    raise KeyError(name)


def main():
    coords, tets, etags = _read_fly(SRC / "cube.fly")
    np.savez_compressed(DST / "cube_mesh.npz", coords=coords, tets=tets, elem_tags=etags)
    shutil.copyfile(SRC / "cube.krn", DST / "cube.krn")
    shutil.copyfile(SRC / "loop" / "cube.p2", DST / "cube.p2")
    shutil.copyfile(SRC / "loop" / "cube.dat", DST / "loop_cube.dat")
    shutil.copyfile(SRC / "hmag" / "cube.csv", DST / "hmag_cube.csv")
    for k in ("vortex", "uniform"):
        shutil.copyfile(SRC / "exani" / f"cube_{k}.csv", DST / f"exani_cube_{k}.csv")
    shutil.copyfile(SRC / "external" / "cube.csv", DST / "external_cube.csv")
    for k in ("anisotropy", "exchange", "hmag", "zeeman", "energy"):
        shutil.copyfile(SRC / "mapping" / f"cube_{k}.csv", DST / f"mapping_cube_{k}.csv")
    lines = (SRC / "loop" / "cube_0001.vtu").read_text().splitlines()
    m = np.array(_vtu_array(lines, "m"), dtype=np.float64).reshape(-1, 3)
    u = np.array(_vtu_array(lines, "u"), dtype=np.float64)
    conn = np.array(_vtu_array(lines, "connectivity"), dtype=np.int64).reshape(-1, 4)
    key = {tuple(t): i for i, t in enumerate(tets.tolist())}
    cell_order = np.array([key[tuple(c)] for c in conn.tolist()], dtype=np.int32)
    np.savez_compressed(DST / "cube_0001_vtu.npz", m=m, u=u, cell_order=cell_order)


if __name__ == "__main__":
    main()
