"""GPU structured (Kuhn-grid stencil) path and mixed precision against the FP64 oracle,
on a small generated 4-grain Voronoi cube (24 nm, 8^3 magnet voxels, graded air box).

FP64: operators to 1e-12 relative, PCG with jax CG semantics, one NCG field step and a
short loop to 1e-9.  Mixed precision (FP32 operators and inner PCG, FP64 dots, energies
and outer residual): energies to 1e-6 relative, M.h to 1e-6 absolute.
"""
import numpy as np
import pytest

from mammos_mumag_amd import generate

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def mesh():
    return generate.build(edge=24.0, n=8, layers=3, grains=4, cone=15.0, seed=7)


@pytest.fixture(scope="module")
def oracle(mesh):
    from oracle.assemble import Model
    return Model(mesh.coords(), mesh.tets(), mesh.elem_tags(), mesh.krn())


@pytest.fixture(scope="module")
def gpu_s(mesh):
    from mammos_mumag_amd import Model
    print("[gpu] structured context", flush=True)
    return Model.from_structured(mesh)


@pytest.fixture(scope="module")
def gpu_g(mesh):
    from mammos_mumag_amd import Model
    print("[gpu] general context on the generated mesh", flush=True)
    return Model.from_structured(mesh, structured=False)


@pytest.fixture(scope="module")
def gpu_mixed(mesh):
    from mammos_mumag_amd import Model
    print("[gpu] mixed-precision structured context (FP32 evaluation kernels)", flush=True)
    return Model.from_structured(mesh, precision="mixed", eval_precision="mixed")


def _diag(model):
    import ctypes
    from mammos_mumag_amd import _lib
    dex = np.empty(3 * model.N)
    dk = np.empty(model.N)
    _lib.check(model._lib.mm_get_diag_host(model._ctx, dex.ctypes.data_as(_lib.c_double_p),
                                           dk.ctypes.data_as(_lib.c_double_p)), "diag")
    return dex, dk


@pytest.mark.parametrize("which", ["gpu_s", "gpu_g"])
def test_operators(which, request, oracle):
    model = request.getfixturevalue(which)
    O = oracle
    assert model.info["structured"] == (1 if which == "gpu_s" else 0)
    assert abs(model.V / O.V - 1) < 1e-14
    assert rel(model.meas, O.meas) < 1e-13
    dex, dk = _diag(model)
    assert rel(dex, O.Dex) < 1e-13
    assert rel(dk, O.diagK) < 1e-12
    rng = np.random.default_rng(11)
    v = rng.normal(size=3 * O.N)
    x = rng.normal(size=O.N)
    assert rel(model.apply_C(model.to_device(v)).cpu().numpy(), O.C @ v) < 1e-12
    assert rel(model.apply_K(model.to_device(x)).cpu().numpy(), O.K @ x) < 1e-12
    b = model.divergence(model.to_device(v)).cpu().numpy()
    assert rel(b, O.Dxyz[0] @ v[0::3] + O.Dxyz[1] @ v[1::3] + O.Dxyz[2] @ v[2::3]) < 1e-12
    g = model.gradient(model.to_device(x)).cpu().numpy()
    assert rel(g, np.stack([O.Gxyz[a] @ x for a in range(3)], 1).ravel()) < 1e-12


def test_pcg_structured(gpu_s, oracle):
    from oracle.solvers import jax_cg
    O = oracle
    m = O.uniform_m([0.0, 0.0, 1.0])
    b = O.Dxyz[2] @ m[2::3]
    x_ref, it_ref = jax_cg(O.K, O.diagK, b, np.zeros(O.N), tol=1e-10)
    x = gpu_s.empty1()
    it, rc = gpu_s.pcg(gpu_s.to_device(b), x, 1e-10)
    assert rc == 0 and abs(it - it_ref) <= 1
    assert rel(x.cpu().numpy(), x_ref) < 1e-9


def test_loop_steps_structured(gpu_s, oracle):
    from mammos_mumag_amd.loop import run_loop
    from oracle.loop import run_loop as oracle_loop
    p = generate.loop_params("cube40_single")
    rec_o, mini = oracle_loop(oracle, p, oracle.uniform_m(p["m"]), max_steps=4)
    rec_g, stats, _ = run_loop(gpu_s, p, gpu_s.uniform_m(p["m"]), max_steps=4, save=False)
    print("[gpu] oracle", mini.stats, "gpu", stats, flush=True)
    for ro, rg in zip(rec_o, rec_g):
        assert ro[0] == rg[0] and ro[1] == rg[1]
        assert abs(rg[2] - ro[2]) < 1e-9
        assert abs(rg[3] / ro[3] - 1) < 1e-9


def test_mixed_precision_errors(gpu_mixed, oracle):
    """Mixed precision keeps the FP64 answer to FP32 operator accuracy; the PCG meets the
    FP64 tolerance on the true residual."""
    from mammos_mumag_amd.loop import run_loop
    from oracle.energies import total_eg
    O = oracle
    p = generate.loop_params("cube40_single")
    rng = np.random.default_rng(5)
    m = O.uniform_m([0.1, 0.2, 0.97]) + 0.05 * rng.normal(size=3 * O.N) * np.repeat(O.meas > 0, 3)
    n = np.linalg.norm(m.reshape(-1, 3), axis=1)
    m = (m.reshape(-1, 3) / np.where(n == 0, 1, n)[:, None]).ravel()
    E_ref, gp_ref, u_ref, _ = total_eg(O, m, 1e-10, 1.5, p["h"], 1, np.zeros(O.N))
    u = gpu_mixed.empty1()
    E, gp, gf, it = gpu_mixed.total_eg(gpu_mixed.to_device(m), 1.5, p["h"], 1e-10, 1, u)
    print(f"[gpu] mixed: dE/E={abs(E / E_ref - 1):.2e} du={rel(u.cpu().numpy(), u_ref):.2e} "
          f"dg={rel(gp.cpu().numpy(), gp_ref):.2e} inner iters={it}", flush=True)
    assert abs(E / E_ref - 1) < 1e-6
    assert rel(u.cpu().numpy(), u_ref) < 1e-5
    assert rel(gp.cpu().numpy(), gp_ref) < 1e-5
    rec, stats, _ = run_loop(gpu_mixed, p, gpu_mixed.uniform_m(p["m"]), max_steps=3, save=False)
    from oracle.loop import run_loop as oracle_loop
    rec_o, _ = oracle_loop(O, p, O.uniform_m(p["m"]), max_steps=3)
    for ro, rg in zip(rec_o, rec):
        assert abs(rg[2] - ro[2]) < 1e-6
        assert abs(rg[3] / ro[3] - 1) < 1e-6


@pytest.fixture(scope="module")
def near_equilibrium(oracle):
    """The oracle's minimised state of the first field step (+1.5 T): its projected
    gradient is ~4e-9, so the truncated CG's stop test (r <= min(0.5, sqrt|g|) |g|) lets
    it run 14 iterations."""
    from oracle.energies import total_eg
    from oracle.loop import Minimizer
    O = oracle
    p = generate.loop_params("cube40_single")
    mini = Minimizer(O, p, O.uniform_m(p["m"]))
    _, x, _ = mini.minimize(O.uniform_m(p["m"]), 1.5)
    _, gp, _, gf = total_eg(O, x, 1e-10, 1.5, p["h"], 1, np.zeros(O.N))
    return gp, gf


@pytest.mark.parametrize("max_iter", [3, 10, 17])
def test_truncated_cg_iteration_caps(gpu_s, gpu_mixed, oracle, near_equilibrium, max_iter):
    """M_inv on the magnet block (k_tcgc_*) with iteration caps below, at and above the
    direction ring's initial 11 slots (17 grows it and, with 14 iterations, uses slots past
    the initial ones), against the oracle's truncated CG: the FP64 structured context with
    the oracle's iteration count to 1e-12, the FP32 form within one iteration to 1e-6."""
    from oracle.solvers import truncated_cg
    from oracle.loop import Minimizer
    O = oracle
    gp_ref, gf_ref = near_equilibrium
    p = dict(generate.loop_params("cube40_single"), precond_iter=max_iter)
    m0 = O.uniform_m(p["m"])
    mini = Minimizer(O, p, m0)
    mini.gF = gf_ref
    z_ref, it_ref = truncated_cg(mini.H, -gp_ref, O.Dex, max_iter=max_iter)
    assert it_ref == min(max_iter, 14)
    for model, tol, dit in ((gpu_s, 1e-12, 0), (gpu_mixed, 1e-6, 1)):
        z, it = model.M_inv(model.to_device(m0), model.to_device(gf_ref), model.to_device(gp_ref), max_iter)
        err = rel(z.cpu().numpy(), z_ref)
        print(f"[gpu] M_inv cap {max_iter} ({model.precision}): iterations {it} (oracle {it_ref}), "
              f"rel {err:.2e}", flush=True)
        assert abs(it - it_ref) <= dit
        assert err < tol


def test_poisson_stencil_chunked_grid():
    """A 66^3-node grid, where the z-march leaves the last block chunks empty: the stencil
    must touch nothing outside the grid and cover every row once (regression)."""
    from mammos_mumag_amd import Model
    from oracle.assemble import Model as OracleModel
    g = generate.build(edge=39.0, n=39, layers=13, grains=2, cone=15.0, seed=3)
    assert len(g.xs) == 66
    model = Model.from_structured(g)
    O = OracleModel(g.coords(), g.tets(), g.elem_tags(), g.krn())
    x = np.random.default_rng(2).normal(size=O.N)
    assert rel(model.apply_K(model.to_device(x)).cpu().numpy(), O.K @ x) < 1e-12


@pytest.mark.parametrize("n,layers,levels", [(8, 3, 1), (39, 13, 4), (125, 0, 6)])
def test_multigrid_vcycle_matches_oracle(n, layers, levels):
    """The GPU V-cycle (FP32) against its FP64 restatement: one level (15^3 nodes), four
    (66^3 -> 40^3 -> 22^3 -> 12^3, the graded air box coarsening last) and the six of the
    64-grain 125^3-voxel benchmark grid (154^3 -> ... -> 9^3), on a fresh context (the
    set-up kernels of mm_multigrid must not race the zeroing of new buffers)."""
    import torch
    from mammos_mumag_amd import Model
    from oracle import mg
    if n == 125:
        g = generate.build_config("voronoi64_250nm", seed=1)
    else:
        g = generate.build(edge=39.0 if n == 39 else 24.0, n=n, layers=layers, grains=2, cone=15.0, seed=3)
    print(f"[gpu] multigrid V-cycle, {len(g.xs)}^3 nodes", flush=True)
    model = Model.from_structured(g, precision="mixed")
    model.multigrid()
    V = mg.VCycle(g.xs, g.ys, g.zs)
    assert len(V.levels) == levels
    assert model.mg_levels() == [(len(a), len(b), len(c)) for a, b, c in V.axes]
    rng = np.random.default_rng(8)
    _, face = mg.operator(g.xs, g.ys, g.zs)
    r = rng.normal(size=model.N)
    r[face] = 0.0
    z = model.mg_apply(torch.tensor(r, dtype=torch.float32, device="cuda")).cpu().numpy().astype(np.float64)
    zr = V.apply(r)
    print(f"[gpu] V-cycle levels {model.mg_levels()}: relative difference {rel(z, zr):.2e}", flush=True)
    assert rel(z, zr) < 1e-5
    model.close()


def test_mixed_multigrid_loop():
    """Mixed precision with the multigrid preconditioner: the loop agrees with plain FP64
    to the mixed form's tolerances, with far fewer PCG iterations than Jacobi."""
    from mammos_mumag_amd import Model, _lib
    g = generate.build(edge=39.0, n=39, layers=13, grains=2, cone=15.0, seed=3)
    p = generate.loop_params("cube40_single")
    out = {}
    for form in ("fp64", "mixed", "mixed+mg"):
        print(f"[gpu] 66^3 loop, {form}", flush=True)
        model = Model.from_structured(g, precision="fp64" if form == "fp64" else "mixed")
        if form == "mixed+mg":
            model.multigrid()
        m = model.to_device(model.uniform_m(p["m"]))
        m_ref, u, gF = m.clone(), model.empty1(), model.empty3()
        st = _lib.Stats()
        hext, rec = p["hstart"], []
        for _ in range(3):
            E = model.minimize(m, hext, p["h"], p, m_ref, u, gF, st)
            rec.append((E, model.compute_mh(m, p["h"])))
            hext -= 0.4
        out[form] = (np.array(rec), st.as_dict())
        model.close()
    print("[gpu] PCG iterations:", {k: v[1]["pcg_iters"] for k, v in out.items()}, flush=True)
    r0 = out["fp64"][0]
    for form in ("mixed", "mixed+mg"):
        r1 = out[form][0]
        assert np.max(np.abs(r1[:, 1] - r0[:, 1])) < 1e-6, form
        assert np.max(np.abs(r1[:, 0] / r0[:, 0] - 1)) < 1e-6, form
    assert out["mixed+mg"][1]["pcg_iters"] < out["mixed"][1]["pcg_iters"] / 10


def test_secant_start_keeps_the_answer():
    """mm_secant_start: the line search's later potential solves start from the Galerkin
    step along the solution's change since its first one.  Plain FP64 (Jacobi) and the
    mixed form with multigrid, each with and without it: the loop's energies and M.h agree
    to within the solve tolerance's effect, with fewer PCG iterations."""
    from mammos_mumag_amd import Model, _lib
    g = generate.build(edge=39.0, n=39, layers=13, grains=2, cone=15.0, seed=3)
    p = generate.loop_params("cube40_single")
    out = {}
    for form in ("fp64", "fp64+secant", "mixed+mg", "mixed+mg+secant"):
        print(f"[gpu] 66^3 loop, {form}", flush=True)
        model = Model.from_structured(g, precision="fp64" if form.startswith("fp64") else "mixed")
        if "mg" in form:
            model.multigrid()
        if form.endswith("secant"):
            model.secant_start(True)
        m = model.to_device(model.uniform_m(p["m"]))
        m_ref, u, gF = m.clone(), model.empty1(), model.empty3()
        st = _lib.Stats()
        hext, rec = p["hstart"], []
        for _ in range(3):
            E = model.minimize(m, hext, p["h"], p, m_ref, u, gF, st)
            rec.append((E, model.compute_mh(m, p["h"])))
            hext -= 0.4
        out[form] = (np.array(rec), st.as_dict())
        model.close()
    for base in ("fp64", "mixed+mg"):
        r0, s0 = out[base]
        r1, s1 = out[base + "+secant"]
        dmh = np.max(np.abs(r1[:, 1] - r0[:, 1]))
        dE = np.max(np.abs(r1[:, 0] / r0[:, 0] - 1))
        print(f"[gpu] {base}: secant |dM.h| {dmh:.2e} |dE/E| {dE:.2e} PCG {s0['pcg_iters']} -> {s1['pcg_iters']} "
              f"NCG {s0['ncg_iters']} -> {s1['ncg_iters']}", flush=True)
        assert dmh < 1e-9 and dE < 1e-11, base
        assert s1["pcg_iters"] < s0["pcg_iters"], base


def test_mixed_multigrid_pcg_limits():
    """mm_pcg in the mixed form with the multigrid preconditioner (its segment target and
    iteration cap are tested on the device between KB and the V-cycle): from zero the true
    residual meets the FP64 path's bound r.r <= tol^2 b.b, with far fewer iterations than
    plain FP64; a cap of 3 iterations returns the warning code after exactly 3; a converged
    start does nothing."""
    import torch
    from mammos_mumag_amd import Model
    from oracle import mg
    g = generate.build(edge=39.0, n=39, layers=13, grains=2, cone=15.0, seed=3)
    A, face = mg.operator(g.xs, g.ys, g.zs)
    rng = np.random.default_rng(5)
    b = rng.normal(size=A.shape[0])
    b[face] = 0.0
    ref = Model.from_structured(g, precision="fp64")
    xr = ref.empty1()
    itr, rc = ref.pcg(ref.to_device(b), xr, 1e-10)
    assert rc == 0
    ref.close()
    model = Model.from_structured(g, precision="mixed")
    model.multigrid()
    x = model.empty1()
    it, rc = model.pcg(model.to_device(b), x, 1e-10)
    xs = x.cpu().numpy()
    res = np.linalg.norm(b - A @ xs) / np.linalg.norm(b)
    print(f"[gpu] mixed+mg pcg: {it} iterations (fp64 Jacobi {itr}), true residual {res:.2e}", flush=True)
    assert rc == 0 and it < itr / 10
    assert res <= 1e-10 * (1 + 1e-6)
    x = model.empty1()
    it, rc = model.pcg(model.to_device(b), x, 1e-14, maxiter=3)
    assert it == 3 and rc == 1
    x = model.to_device(xs)
    x0 = x.clone()
    it, rc = model.pcg(model.to_device(b), x, 1e-6)
    assert it == 0 and rc == 0 and torch.equal(x, x0)
    model.close()


def test_multigrid_setup_toggle():
    """mm_multigrid(ctx, 0) restores the Jacobi form with its own reliable-update interval
    (the same iteration count as a context that never had the multigrid); a second set-up
    with the same bound re-enables the built levels without allocating; another bound is
    refused."""
    from mammos_mumag_amd import Model, _lib
    g = generate.build(edge=39.0, n=39, layers=13, grains=2, cone=15.0, seed=3)
    rng = np.random.default_rng(8)
    fresh = Model.from_structured(g, precision="mixed")
    b = fresh.divergence(fresh.to_device(fresh.state_m([0, 0, 1], 1)))
    b = b + 0.01 * fresh.to_device(rng.normal(size=fresh.N)) * fresh.to_device((fresh.meas > 0).astype(float))
    x = fresh.empty1()
    it_fresh, _ = fresh.pcg(b, x, 1e-10)
    fresh.close()
    model = Model.from_structured(g, precision="mixed")
    model.multigrid(8)
    levels = model.mg_levels()
    it_mg, _ = model.pcg(b, model.empty1(), 1e-10)
    model.multigrid(0)
    assert model.mg_levels() == []
    it_back, _ = model.pcg(b, model.empty1(), 1e-10)
    print(f"[gpu] Jacobi {it_fresh} iterations fresh, {it_back} after multigrid and back; multigrid {it_mg}",
          flush=True)
    assert it_back == it_fresh
    import ctypes
    info = _lib.Info()
    _lib.check(model._lib.mm_ctx_info(model._ctx, ctypes.byref(info)), "info")
    before = info.device_bytes
    model.multigrid(8)
    assert model.mg_levels() == levels
    assert model.pcg(b, model.empty1(), 1e-10)[0] == it_mg
    _lib.check(model._lib.mm_ctx_info(model._ctx, ctypes.byref(info)), "info")
    assert info.device_bytes == before
    with pytest.raises(_lib.MumagError):
        model.multigrid(6)
    model.close()
