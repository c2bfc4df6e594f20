"""The drop-in entry points through ``Simulation`` (the reference's file contract), on the
reference's test mesh, against the reference's own expected outputs — the same checks as
the reference's tests/test_{exani,external,hmag,loop,mapping}.py (np.allclose on the CSV
values, the .dat loop and the VTU ``m``), plus tighter bounds where FP64 allows.

Each entry point runs in a child process (``python -m mammos_mumag_amd.scripts.<s>``)
that builds its own device context.
"""
import numpy as np
import pytest

from conftest import read_csv_values

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim(cube_dir):
    from mammos_mumag_amd.simulation import Simulation
    return Simulation(mesh_filepath=cube_dir / "cube.fly", materials_filepath=cube_dir / "cube.krn",
                      parameters_filepath=cube_dir / "cube.p2")


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.maximum(np.abs(b), 1e-300)))


def test_exani(sim, tmp_path, golden):
    print("[gpu] run_exani", flush=True)
    sim.run_exani(outdir=tmp_path)
    for case in ("vortex", "uniform"):
        ref = read_csv_values(golden / f"exani_cube_{case}.csv")
        out = read_csv_values(tmp_path / f"out_{case}.csv")
        assert np.allclose(ref, out)
        assert _rel(out, ref) < 1e-9, case
    assert (tmp_path / "info.json").exists()


def test_external(sim, tmp_path, golden):
    print("[gpu] run_external", flush=True)
    sim.run_external(outdir=tmp_path)
    ref = read_csv_values(golden / "external_cube.csv")
    out = read_csv_values(tmp_path / "out.csv")
    assert np.allclose(ref, out)
    assert _rel(out, ref) < 1e-12


def test_hmag(sim, tmp_path, golden, oracle_cube):
    print("[gpu] run_hmag", flush=True)
    sim.run_hmag(outdir=tmp_path, name="cube")
    ref = read_csv_values(golden / "hmag_cube.csv")
    out = read_csv_values(tmp_path / "cube.csv")
    assert np.allclose(ref, out)
    assert _rel(out, ref) < 1e-9
    # the VTU's fields (hmag.py:153-165) against the oracle: U solves K U = D m, h_nodes =
    # -V (G U) / meas (meas 0 -> 1), cell h = -grad U; written with 7 significant digits
    from mammos_mumag_amd.fileio import read_vtu_array, vtu_cell_order
    from scipy.sparse.linalg import spsolve
    O = oracle_cube
    m = O.uniform_m([0.0, 0.0, 1.0])
    b = sum(O.Dxyz[a] @ m[a::3] for a in range(3))
    u_ref = spsolve(O.K.tocsc(), b)
    g_ref = np.stack([O.Gxyz[a] @ u_ref for a in range(3)], axis=1)
    meas = np.where(O.meas == 0, 1.0, O.meas)
    hn_ref = -O.V * g_ref / meas[:, None]
    X = O.coords[O.tets]
    du = u_ref[O.tets[:, 1:]] - u_ref[O.tets[:, 0:1]]
    hc_ref = -np.linalg.solve(X[:, 1:, :] - X[:, 0:1, :], du[..., None])[..., 0]
    vtu = tmp_path / "cube_hmag.vtu"
    u = read_vtu_array(vtu, "U")
    hn = read_vtu_array(vtu, "h_nodes")
    hc = read_vtu_array(vtu, "h")
    mv = read_vtu_array(vtu, "m")
    order = vtu_cell_order(O.tets)
    assert np.max(np.abs(u - u_ref)) <= 2e-6 * np.max(np.abs(u_ref))
    assert np.max(np.abs(hn - hn_ref)) <= 2e-6 * np.max(np.abs(hn_ref))
    assert np.max(np.abs(hc - hc_ref[order])) <= 2e-6 * np.max(np.abs(hc_ref))
    assert np.array_equal(mv, m.reshape(-1, 3))


def test_loop(sim, tmp_path, golden):
    print("[gpu] run_loop", flush=True)
    sim.run_loop(outdir=tmp_path, name="cube")
    data_loop = np.loadtxt(golden / "loop_cube.dat")
    sim_loop = np.loadtxt(tmp_path / "cube.dat")
    assert sim_loop.shape == data_loop.shape
    assert np.allclose(data_loop, sim_loop)
    assert np.max(np.abs(sim_loop[:, 2] - data_loop[:, 2])) < 1e-6      # M.h
    vtus = sorted(p.name for p in tmp_path.iterdir() if p.suffix == ".vtu")
    assert "cube_0001.vtu" in vtus
    from mammos_mumag_amd.fileio import read_vtu_array
    ref = np.load(golden / "cube_0001_vtu.npz")
    m = read_vtu_array(tmp_path / "cube_0001.vtu", "m")
    assert np.allclose(ref["m"].reshape(m.shape), m)


@pytest.mark.parametrize("amg", ["1", "0"])
def test_loop_mixed(sim, tmp_path, golden, monkeypatch, amg):
    """The loop script in the mixed form (MAMMOS_MUMAG_PRECISION=mixed), with the AMG
    preconditioner (the default of that form) and with Jacobi: the reference's records to
    its own criterion (np.allclose) and M.h to 1e-6."""
    monkeypatch.setenv("MAMMOS_MUMAG_PRECISION", "mixed")
    monkeypatch.setenv("MAMMOS_MUMAG_AMG", amg)
    print(f"[gpu] run_loop mixed, MAMMOS_MUMAG_AMG={amg}", flush=True)
    sim.run_loop(outdir=tmp_path, name="cube")
    data_loop = np.loadtxt(golden / "loop_cube.dat")
    sim_loop = np.loadtxt(tmp_path / "cube.dat")
    assert sim_loop.shape == data_loop.shape
    assert np.array_equal(sim_loop[:, :2], data_loop[:, :2])
    assert np.allclose(data_loop, sim_loop)
    assert np.max(np.abs(sim_loop[:, 2] - data_loop[:, 2])) < 1e-6      # M.h


def test_mapping(sim, tmp_path, golden):
    """mapping.py evaluates in float32 (scripts/mapping.py:22-118): the fixtures hold
    float32 values, compared as the reference's test_mapping does."""
    print("[gpu] run_mapping", flush=True)
    sim.run_mapping(outdir=tmp_path, name="cube")
    for part in ("zeeman", "exchange", "anisotropy", "hmag", "energy"):
        ref = read_csv_values(golden / f"mapping_cube_{part}.csv")
        out = read_csv_values(tmp_path / f"cube_{part}.csv")
        assert np.allclose(ref, out, rtol=1e-5), part
