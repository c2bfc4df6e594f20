"""File formats of the drop-in contract.

* ``.fly`` finley text mesh (read by the reference through ``scripts/mesh.py:12``,
  written by ``tofly.py:259-337``): ``3D-nodes N`` + ``id gDOF tag x y z``,
  ``Tet4 NE`` + ``id tag n1 n2 n3 n4`` (1-based), then empty ``Tri3``,
  ``Tri3_Contact``, ``Point1`` sections and ``Tags``.
* ``.krn``: line i = i-th sorted element tag, ``theta phi K1 K2 Js A ...``
  (``scripts/materials.py:22-34``).
* ``.p2``: INI read with the defaults of ``scripts/tools.py:64-120``.
* ``.dat``: ``int(counter) hext mh E/mu0`` (``scripts/tools.py:92-96``).
* VTU: ASCII UnstructuredGrid as written by weipa's saveVTK
  (``scripts/escript_tools.py:26-31``): points in node order, cells sorted
  stably by their smallest node index, ``%e`` values, ``tags`` cell data.
* CSV: ``title`` line then ``name,value,explanation`` rows (``scripts/hmag.py:141-152``).
"""
import configparser
import inspect
import math

import numpy as np
import scipy.constants


def mu0_output():
    """Outputs divide by scipy's mu_0 (``scripts/tools.py:98-99``)."""
    return scipy.constants.mu_0


MU0_SCRIPT = 4.0e-7 * math.pi  # material scaling, scripts/materials.py:21


def read_fly(path):
    """Return coords f64[N,3], tets i32[NE,4] (0-based), elem_tags i32[NE]."""
    with open(path) as f:
        lines = f.read().splitlines()
    hdr = lines[1].split()
    if hdr[0] != "3D-nodes":
        raise ValueError(f"{path}: expected '3D-nodes', got {lines[1]!r}")
    n = int(hdr[1])
    nodes = np.array(" ".join(lines[2:2 + n]).split(), dtype=np.float64).reshape(n, 6)
    ids = nodes[:, 0].astype(np.int64)
    if not np.all(np.diff(ids) > 0):
        order = np.argsort(ids, kind="stable")
        nodes, ids = nodes[order], ids[order]
    thdr = lines[2 + n].split()
    if thdr[0] != "Tet4":
        raise ValueError(f"{path}: expected 'Tet4', got {lines[2 + n]!r}")
    ne = int(thdr[1])
    el = np.array(" ".join(lines[3 + n:3 + n + ne]).split(), dtype=np.int64).reshape(ne, 6)
    id2idx = np.full(int(ids.max()) + 1, -1, dtype=np.int64)
    id2idx[ids] = np.arange(n)
    tets = id2idx[el[:, 2:6]]
    if (tets < 0).any():
        raise ValueError(f"{path}: element refers to an unknown node id")
    return nodes[:, 3:6].copy(), tets.astype(np.int32), el[:, 1].astype(np.int32)


def write_fly(path, coords, tets, elem_tags, name="Mesh"):
    n, ne = len(coords), len(tets)
    with open(path, "w") as f:
        f.write(f"{name}\n3D-nodes {n}\n")
        for i in range(0, n, 200000):
            blk = coords[i:i + 200000]
            f.write("".join(f"{i + k + 1} {i + k + 1} 0 {x:.16E} {y:.16E} {z:.16E}\n"
                            for k, (x, y, z) in enumerate(blk)))
        f.write(f"Tet4 {ne}\n")
        t1 = tets + 1
        for i in range(0, ne, 200000):
            blk = t1[i:i + 200000]
            tg = elem_tags[i:i + 200000]
            f.write("".join(f"{i + k + 1} {tg[k]} {a} {b} {c} {d}\n"
                            for k, (a, b, c, d) in enumerate(blk.tolist())))
        f.write("Tri3 0\nTri3_Contact 0\nPoint1 0\nTags\n")


def read_krn(path, tags):
    """Per sorted tag: dict(theta, phi, K1, K2, Js, A)."""
    out = {}
    with open(path) as f:
        lines = [ln.split() for ln in f.read().splitlines() if ln.strip()]
    if len(lines) < len(tags):
        raise ValueError(f"{path}: {len(lines)} material lines for {len(tags)} tags")
    for tag, line in zip(tags, lines):
        out[int(tag)] = dict(theta=float(line[0]), phi=float(line[1]), K1=float(line[2]),
                             K2=float(line[3]), Js=float(line[4]), A=float(line[5]))
    return out


def write_krn(path, rows):
    """rows: iterable of (theta, phi, K1, K2, Js, A)."""
    with open(path, "w") as f:
        for r in rows:
            f.write(" ".join(repr(float(v)) for v in r) + " 1.0 0.0\n")


def _normalize(v):
    a, b, c = v
    s = math.sqrt(a * a + b * b + c * c)
    return [a, b, c] if s == 0.0 else [a / s, b / s, c / s]


def read_p2(path):
    """``scripts/tools.py:64-120`` read_params, returned as a dict."""
    config = configparser.ConfigParser({
        "mstep": 1.0, "state": "mxyz", "mfinal": -0.8, "hmag_on": 1, "tol_fun": 1e-10,
        "tol_hmag_factor": 1.0, "precond_iter": 10, "iter_max": 1000, "verbose": 1,
    })
    if not config.read(path):
        raise FileNotFoundError(path)
    ini, field, mini = config["initial state"], config["field"], config["minimizer"]
    state = ini["state"].lower()
    tol_fun = float(mini["tol_fun"])
    return dict(
        m=_normalize([float(ini["mx"]), float(ini["my"]), float(ini["mz"])]),
        state_id={"flower": 1, "vortex": 2, "twisted": 3, "random": 4}.get(state, 0),
        mstep=float(field["mstep"]), mfinal=float(field["mfinal"]),
        h=_normalize([float(field["hx"]), float(field["hy"]), float(field["hz"])]),
        hstart=float(field["hstart"]), hfinal=float(field["hfinal"]), hstep=float(field["hstep"]),
        hmag_on=int(mini["hmag_on"]), tol_fun=tol_fun,
        tol_u=tol_fun * float(mini["tol_hmag_factor"]),
        precond_iter=int(mini["precond_iter"]), iter_max=int(mini["iter_max"]),
        verbose=int(mini["verbose"]),
    )


def write_p2(path, p):
    """Writer for the ``.p2`` template fields (``pkg/templates/p2.jinja``)."""
    with open(path, "w") as f:
        f.write(inspect.cleandoc(f"""
            [mesh]
            size = {p.get('size', 1e-9)}
            scale = {p.get('scale', 0.0)}

            [initial state]
            state = {p.get('state', 'mxyz')}
            mx = {p['m'][0]}
            my = {p['m'][1]}
            mz = {p['m'][2]}

            [field]
            hstart = {p['hstart']}
            hfinal = {p['hfinal']}
            hstep = {p['hstep']}
            hx = {p['h'][0]}
            hy = {p['h'][1]}
            hz = {p['h'][2]}
            mstep = {p.get('mstep', 1.0)}
            mfinal = {p.get('mfinal', -0.8)}

            [minimizer]
            hmag_on = {p.get('hmag_on', 1)}
            iter_max = {p.get('iter_max', 1000)}
            tol_fun = {p.get('tol_fun', 1e-10)}
            tol_hmag_factor = {p.get('tol_hmag_factor', 1.0)}
            precond_iter = {p.get('precond_iter', 10)}
            """) + "\n")


def write_dat(path, rec):
    mu0 = mu0_output()
    with open(path, "w") as f:
        for c, hext, mh, e in rec:
            f.write(f"{int(c)} {float(hext)!r} {float(mh)!r} {float(e) / mu0!r}\n")


def write_csv(path, title, rows):
    """rows: list of (name, value, explanation)."""
    with open(path, "w") as f:
        f.write(title + "\nname,value,explanation\n")
        f.write("\n".join(f"{n},{float(v)!r},{e}" for n, v, e in rows))


def _fmt(a):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    return "\n".join(" ".join("%e" % v for v in row) for row in a.tolist())


def vtu_cell_order(tets):
    return np.argsort(tets.min(axis=1), kind="stable")


def write_vtu(path, coords, tets, elem_tags, point_data, cell_data=None):
    order = vtu_cell_order(tets)
    t = tets[order]
    out = ['<?xml version="1.0"?>', '<VTKFile type="UnstructuredGrid" version="0.1">',
           "<UnstructuredGrid>", "<FieldData>",
           '<DataArray Name="TIME" type="Float64" format="ascii" NumberOfTuples="1">',
           "0.000000e+00", "</DataArray>", "</FieldData>", "<FieldData>",
           '<DataArray Name="CYCLE" type="Int32" format="ascii" NumberOfTuples="1">', "0",
           "</DataArray>", "</FieldData>",
           f'<Piece NumberOfPoints="{len(coords)}" NumberOfCells="{len(t)}">', "<Points>",
           '<DataArray NumberOfComponents="3" type="Float64" format="ascii">', _fmt(coords),
           "</DataArray>", "</Points>", "<Cells>",
           '<DataArray Name="connectivity" type="Int32" format="ascii">',
           "\n".join(f"{a} {b} {c} {d}" for a, b, c, d in t.tolist()), "</DataArray>",
           '<DataArray Name="offsets" type="Int32" format="ascii">',
           "\n".join(str(4 * (i + 1)) for i in range(len(t))), "</DataArray>",
           '<DataArray Name="types" type="UInt8" format="ascii">', "\n".join(["10"] * len(t)),
           "</DataArray>", "</Cells>", "<PointData>"]
    for name, arr in point_data.items():
        arr = np.asarray(arr, dtype=np.float64)
        nc = 1 if arr.ndim == 1 else arr.shape[1]
        out += [f'<DataArray Name="{name}" type="Float64" NumberOfComponents="{nc}" format="ascii">',
                _fmt(arr), "</DataArray>"]
    out += ["</PointData>", "<CellData>"]
    cells = {"tags": np.asarray(elem_tags, dtype=np.float64)}
    if cell_data:
        cells.update(cell_data)
    for name, arr in cells.items():
        arr = np.asarray(arr, dtype=np.float64)[order]
        nc = 1 if arr.ndim == 1 else arr.shape[1]
        out += [f'<DataArray Name="{name}" type="Float64" NumberOfComponents="{nc}" format="ascii">',
                _fmt(arr), "</DataArray>"]
    out += ["</CellData>", "</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def read_vtu_array(path, name):
    with open(path) as f:
        lines = f.read().splitlines()
    for i, line in enumerate(lines):
        if line.startswith("<DataArray") and f'Name="{name}"' in line:
            j = i + 1
            while not lines[j].startswith("</DataArray>"):
                j += 1
            nc = int(line.split('NumberOfComponents="')[1].split('"')[0]) if "NumberOfComponents" in line else 1
            a = np.array(" ".join(lines[i + 1:j]).split(), dtype=np.float64)
            return a.reshape(-1, nc) if nc > 1 else a
    raise KeyError(name)
