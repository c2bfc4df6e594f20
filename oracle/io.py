"""Oracle file I/O: .fly / .krn / .p2 readers, .dat and weipa-style VTU writers.

Test infrastructure only (see ``oracle/__init__.py``).

Reference behaviour restated here:

* ``.fly`` (finley text mesh) as read by ``scripts/mesh.py:12`` (ReadMesh):
  ``3D-nodes N`` then ``id gDOF tag x y z``; ``Tet4 NE`` then
  ``id tag n1 n2 n3 n4`` (1-based node ids).  Nodes are kept in file order.
* ``.krn``: line i belongs to the i-th sorted element tag
  (``scripts/materials.py:22-34``); columns ``theta phi K1 K2 Js A``.
* ``.p2``: ``scripts/tools.py:64-120`` read_params, same defaults.
* ``.dat``: ``scripts/tools.py:92-96`` write_mh.
* VTU: ``scripts/escript_tools.py:26-31`` (weipa saveVTK): points in
  node order, cells sorted stably by their smallest node index, ASCII ``%e``.
"""
import configparser
import math

import numpy as np
import scipy.constants

MU0_SCRIPT = 4.0e-7 * math.pi  # scripts/materials.py:21


def mu0_output():
    """scripts/tools.py:98-99 -- outputs divide by scipy's CODATA mu_0."""
    return scipy.constants.mu_0


def read_fly(path):
    """Return (name, coords f64[N,3], node_tags i32[N], tets i32[NE,4] 0-based, elem_tags i32[NE])."""
    with open(path) as f:
        lines = f.read().splitlines()
    name = lines[0].strip()
    hdr = lines[1].split()
    assert hdr[0] == "3D-nodes"
    n = int(hdr[1])
    nodes = np.array(" ".join(lines[2:2 + n]).split(), dtype=np.float64).reshape(n, 6)
    ids = nodes[:, 0].astype(np.int64)
    order = np.argsort(ids, kind="stable")
    if not np.array_equal(ids, np.sort(ids)):
        nodes = nodes[order]
        ids = ids[order]
    coords = nodes[:, 3:6].copy()
    node_tags = nodes[:, 2].astype(np.int32)
    thdr = lines[2 + n].split()
    assert thdr[0] == "Tet4"
    ne = int(thdr[1])
    el = np.array(" ".join(lines[3 + n:3 + n + ne]).split(), dtype=np.int64).reshape(ne, 6)
    id2idx = np.full(ids.max() + 1, -1, dtype=np.int64)
    id2idx[ids] = np.arange(n)
    tets = id2idx[el[:, 2:6]].astype(np.int32)
    assert (tets >= 0).all()
    return name, coords, node_tags, tets, el[:, 1].astype(np.int32)


def read_krn(path, tags):
    """Per sorted tag: dict(theta, phi, K1, Js, A).  scripts/materials.py:22-35."""
    out = {}
    with open(path) as f:
        for tag in tags:
            line = f.readline().split()
            out[int(tag)] = dict(theta=float(line[0]), phi=float(line[1]), K1=float(line[2]),
                                 Js=float(line[4]), A=float(line[5]))
    return out


def _normalize(v):
    """scripts/tools.py:101-109."""
    a, b, c = v
    s = math.sqrt(a * a + b * b + c * c)
    if s == 0.0:
        return [a, b, c]
    return [a / s, b / s, c / s]


def read_p2(path):
    """scripts/tools.py:64-120 read_params -> dict with the same tuple meaning."""
    config = configparser.ConfigParser({
        "mstep": 1.0, "state": "mxyz", "mfinal": -0.8, "hmag_on": 1, "tol_fun": 1e-10,
        "tol_hmag_factor": 1.0, "precond_iter": 10, "iter_max": 1000, "verbose": 1,
    })
    config.read(path)
    ini, field, mini = config["initial state"], config["field"], config["minimizer"]
    m = _normalize([float(ini["mx"]), float(ini["my"]), float(ini["mz"])])
    state = ini["state"].lower()
    state_id = {"flower": 1, "vortex": 2, "twisted": 3, "random": 4}.get(state, 0)
    h = _normalize([float(field["hx"]), float(field["hy"]), float(field["hz"])])
    tol_fun = float(mini["tol_fun"])
    return dict(
        m=m, mstep=float(field["mstep"]), mfinal=float(field["mfinal"]), state_id=state_id,
        h=h, hstart=float(field["hstart"]), hfinal=float(field["hfinal"]),
        hstep=float(field["hstep"]),
        hmag_on=int(mini["hmag_on"]),  # read from [minimizer] only (scripts/tools.py:106)
        tol_u=tol_fun * float(mini["tol_hmag_factor"]), tol_fun=tol_fun,
        precond_iter=int(mini["precond_iter"]), iter_max=int(mini["iter_max"]),
    )


def write_dat(path, rec):
    """scripts/tools.py:92-96: ``int(counter) hext mh E/mu0`` per field step."""
    mu0 = mu0_output()
    with open(path, "w") as f:
        for c, hext, mh, e in rec:
            f.write(f"{int(c)} {float(hext)!r} {float(mh)!r} {float(e) / mu0!r}\n")


def _fmt_rows(a):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    return "\n".join(" ".join("%e" % v for v in row) for row in a)


def write_vtu(path, coords, tets, elem_tags, point_data):
    """weipa-style ASCII VTU (scripts/escript_tools.py:26-31)."""
    order = np.argsort(tets.min(axis=1), kind="stable")
    t = tets[order]
    parts = ['<?xml version="1.0"?>',
             '<VTKFile type="UnstructuredGrid" version="0.1">', "<UnstructuredGrid>",
             "<FieldData>",
             '<DataArray Name="TIME" type="Float64" format="ascii" NumberOfTuples="1">',
             "0.000000e+00", "</DataArray>", "</FieldData>", "<FieldData>",
             '<DataArray Name="CYCLE" type="Int32" format="ascii" NumberOfTuples="1">',
             "0", "</DataArray>", "</FieldData>",
             f'<Piece NumberOfPoints="{len(coords)}" NumberOfCells="{len(t)}">',
             "<Points>", '<DataArray NumberOfComponents="3" type="Float64" format="ascii">',
             _fmt_rows(coords), "</DataArray>", "</Points>", "<Cells>",
             '<DataArray Name="connectivity" type="Int32" format="ascii">',
             "\n".join(" ".join(str(int(v)) for v in row) for row in t), "</DataArray>",
             '<DataArray Name="offsets" type="Int32" format="ascii">',
             "\n".join(str(4 * (i + 1)) for i in range(len(t))), "</DataArray>",
             '<DataArray Name="types" type="UInt8" format="ascii">',
             "\n".join("10" for _ in range(len(t))), "</DataArray>", "</Cells>", "<PointData>"]
    for name, arr in point_data.items():
        arr = np.asarray(arr, dtype=np.float64)
        nc = 1 if arr.ndim == 1 else arr.shape[1]
        parts += [f'<DataArray Name="{name}" type="Float64" NumberOfComponents="{nc}" format="ascii">',
                  _fmt_rows(arr), "</DataArray>"]
    parts += ["</PointData>", "<CellData>",
              '<DataArray Name="tags" type="Float64" NumberOfComponents="1" format="ascii">',
              _fmt_rows(elem_tags[order].astype(np.float64)), "</DataArray>", "</CellData>",
              "</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
    with open(path, "w") as f:
        f.write("\n".join(parts) + "\n")


def read_vtu_point_data(path, name):
    """Minimal reader for the ASCII VTU written above (and by weipa)."""
    with open(path) as f:
        lines = f.read().splitlines()
    for i, line in enumerate(lines):
        if line.startswith("<DataArray") and f'Name="{name}"' in line:
            j = i + 1
            vals = []
            while not lines[j].startswith("</DataArray>"):
                vals.append(lines[j])
                j += 1
            nc = 1
            if 'NumberOfComponents="' in line:
                nc = int(line.split('NumberOfComponents="')[1].split('"')[0])
            arr = np.array(" ".join(vals).split(), dtype=np.float64)
            return arr.reshape(-1, nc) if nc > 1 else arr
    raise KeyError(name)
