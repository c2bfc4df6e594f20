"""The GPU path (general kernels) on generated spherical-shell meshes: the reference's
shell branch (scripts/matrix.py:49-70, the Imhoff tensor of trafo.py:24-45) above the
9,698 nodes of its test mesh.

* 4-grain cube, 42,875 nodes: set-up (R, Rinf, Dirichlet set) and the operators K, C, D,
  G against the oracle assembled on the same mesh (1e-12 relative).
* The same mesh through a reversal loop (0 T to -4 T, reversal at -4 T) against the
  oracle's records (tests/golden/sphere_shell.json): the same records and reversal, M.h
  within 1e-6 and E within 1e-9 relative (or 3x what rounding alone moves in the oracle
  when its initial state moves by 1e-15: its NCG converges slowly, 204-216 iterations).
* The magnetostatic energy of the uniformly magnetised cube on meshes of 42,875, 103,823
  and 704,969 nodes (4.1 M tets) against the oracle's (1e-10 relative), the PCG iteration
  count within 2 % of the oracle's, and within 2.5 % of the closed form Js^2 / 6.
"""
import json

import numpy as np
import pytest

from mammos_mumag_amd import generate

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def ref(golden):
    return json.loads((golden / "sphere_shell.json").read_text())


@pytest.fixture(scope="module")
def loop_mesh(ref):
    return generate.build_sphere(**ref["loop_mesh"])


def test_setup_and_operators(loop_mesh):
    import oracle.assemble as oa
    from mammos_mumag_amd import Model
    print("[gpu] sphere mesh 42,875 nodes: operators", flush=True)
    S = loop_mesh
    O = oa.Model(S.coords(), S.tets(), S.elem_tags(), S.krn())
    gpu = Model.from_structured(S)
    assert not gpu.bnd["airbox"] and not O.airbox
    assert gpu.bnd["R"] == O.R and gpu.bnd["Rinf"] == O.Rinf
    assert np.array_equal(gpu.bnd["dirichlet"], O.dirichlet)
    assert abs(gpu.V / O.V - 1) < 1e-14
    rng = np.random.default_rng(5)
    v = rng.normal(size=3 * O.N)
    x = rng.normal(size=O.N)
    assert rel(gpu.apply_K(gpu.to_device(x)).cpu().numpy(), O.K @ x) < 1e-12
    assert rel(gpu.apply_C(gpu.to_device(v)).cpu().numpy(), O.C @ v) < 1e-12
    b = gpu.divergence(gpu.to_device(v)).cpu().numpy()
    assert rel(b, O.Dxyz[0] @ v[0::3] + O.Dxyz[1] @ v[1::3] + O.Dxyz[2] @ v[2::3]) < 1e-12
    g = gpu.gradient(gpu.to_device(x)).cpu().numpy()
    assert rel(g, np.stack([O.Gxyz[a] @ x for a in range(3)], axis=1).ravel()) < 1e-12
    gpu.close()


def test_reversal_loop_matches_oracle(loop_mesh, ref):
    from mammos_mumag_amd import Model
    from mammos_mumag_amd.loop import run_loop
    print("[gpu] sphere mesh 42,875 nodes: reversal loop", flush=True)
    gpu = Model.from_structured(loop_mesh)
    params = ref["loop_params"]
    steps = []
    rec, stats, _ = run_loop(gpu, params, gpu.uniform_m(params["m"]), save=False,
                             on_step=lambda i, h, mh, E, st, dt: steps.append(dict(st)))
    gpu.close()
    rec = np.array(rec)
    o = ref["loop"]
    oref, sp = np.array(o["records"]), o["spread"]
    dmh = float(np.max(np.abs(rec[:, 2] - oref[:, 2])))
    dE = float(np.max(np.abs(rec[:, 3] / oref[:, 3] - 1)))
    print(f"[gpu] counters {steps[-1]}; |dM.h| {dmh:.2e} (spread {sp['mh']:.1e}), dE/E {dE:.2e} "
          f"(spread {sp['E_rel']:.1e})", flush=True)
    assert len(rec) in sp["n_records"]
    assert np.array_equal(rec[:, :2], oref[:, :2])
    assert rec[-2, 2] > 0 > rec[-1, 2]
    last = steps[-1]
    for k in ("backtracks", "restarts_g", "restarts_z"):
        assert last[k] == sp[k][0] == sp[k][1], k
    for k, rk in (("ncg_iters", "ncg_iters"), ("fcalls", "calls"), ("tcg_iters", "tcg_iters")):
        lo, hi = sp[rk]
        assert 0.95 * lo <= last[k] <= 1.05 * hi, f"{k}: {last[k]} outside [{lo}, {hi}] +- 5%"
    assert dmh <= max(1e-6, 3 * sp["mh"])
    assert dE <= max(1e-9, 3 * sp["E_rel"])


@pytest.mark.parametrize("key", ["16", "24", "60"])
def test_uniform_demag_energy(key, ref):
    from mammos_mumag_amd import Model
    r = ref["hmag_uniform"][key]
    print(f"[gpu] sphere mesh {r['N']} nodes / {r['NE']} tets: uniform demagnetising energy", flush=True)
    mesh = generate.build_sphere(**r["mesh"])
    gpu = Model.from_structured(mesh)
    assert gpu.N == r["N"] and gpu.NE == r["NE"]
    assert gpu.bnd["R"] == pytest.approx(r["R"], rel=1e-15)
    assert gpu.bnd["Rinf"] == pytest.approx(r["Rinf"], rel=1e-15)
    m = gpu.to_device(gpu.uniform_m([0.0, 0.0, 1.0]))
    E, _, it = gpu.hmag_eg(m, gpu.empty1(), 1e-10)
    gpu.close()
    print(f"[gpu] E {E!r} (oracle {r['E']!r}), PCG {it} (oracle {r['pcg_iters']}), "
          f"E / (Js^2/6) - 1 = {E / ref['analytic'] - 1:.4f}", flush=True)
    assert abs(E / r["E"] - 1) < 1e-10
    assert abs(it - r["pcg_iters"]) <= max(2, 0.02 * r["pcg_iters"])
    assert -0.025 < E / ref["analytic"] - 1 < 0


def test_amg_on_shell_meshes(loop_mesh, ref):
    """The mixed form with the smoothed-aggregation AMG on the shell meshes, whose Imhoff
    tensor makes K anisotropic in the shell: the uniform demagnetising energy of the
    704,969-node mesh against the oracle's (1e-9 relative) with a fraction of Jacobi's PCG
    iterations, and the 42,875-node reversal loop with the oracle's records and reversal
    (M.h within 1e-6 and E within 1e-9 relative, as the FP64 loop above)."""
    from mammos_mumag_amd import Model
    from mammos_mumag_amd.loop import run_loop
    r = ref["hmag_uniform"]["60"]
    print(f"[gpu] sphere mesh {r['N']} nodes: uniform demagnetising energy, mixed + AMG", flush=True)
    gpu = Model.from_structured(generate.build_sphere(**r["mesh"]), precision="mixed")
    gpu.amg()
    levels = gpu.amg_levels()
    m = gpu.to_device(gpu.uniform_m([0.0, 0.0, 1.0]))
    E, _, it = gpu.hmag_eg(m, gpu.empty1(), 1e-10)
    gpu.close()
    print(f"[gpu] AMG levels {levels}: E {E!r} (oracle {r['E']!r}), PCG {it} (oracle, Jacobi: "
          f"{r['pcg_iters']})", flush=True)
    assert abs(E / r["E"] - 1) < 1e-9
    assert it < r["pcg_iters"] / 5

    print("[gpu] sphere mesh 42,875 nodes: reversal loop, mixed + AMG", flush=True)
    gpu = Model.from_structured(loop_mesh, precision="mixed")
    gpu.amg()
    params = ref["loop_params"]
    rec, stats, _ = run_loop(gpu, params, gpu.uniform_m(params["m"]), save=False)
    gpu.close()
    rec = np.array(rec)
    o = ref["loop"]
    oref, sp = np.array(o["records"]), o["spread"]
    dmh = float(np.max(np.abs(rec[:, 2] - oref[:, 2])))
    dE = float(np.max(np.abs(rec[:, 3] / oref[:, 3] - 1)))
    print(f"[gpu] |dM.h| {dmh:.2e}, dE/E {dE:.2e}, PCG iterations {stats['pcg_iters']}", flush=True)
    assert len(rec) in sp["n_records"]
    assert np.array_equal(rec[:, :2], oref[:, :2])
    assert rec[-2, 2] > 0 > rec[-1, 2]
    assert dmh <= max(1e-6, 3 * sp["mh"])
    assert dE <= max(1e-9, 3 * sp["E_rel"])
