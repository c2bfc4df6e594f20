"""GPU smoothed-aggregation AMG of the general (unstructured) path (amg.cpp set-up,
amg.hip V-cycle) against its FP64 restatement (oracle/amg.py), and the mixed form's loop
with it against plain FP64 on the general kernels."""
import numpy as np
import pytest

from mammos_mumag_amd import generate

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def context_K(model):
    """The context's own K (FP64), column by column through mm_apply_K: the set-up's
    aggregation breaks ties between equal couplings by index, so the restatement must see
    the same matrix bits (the oracle's assembly agrees with it only to ~1e-14)."""
    import scipy.sparse as sp
    import torch
    rows, cols, vals = [], [], []
    e = torch.zeros(model.N, dtype=torch.float64, device="cuda")
    for j in range(model.N):
        e.zero_()
        e[j] = 1.0
        c = model.apply_K(e).cpu().numpy()
        nz = np.flatnonzero(c)
        rows.append(nz)
        cols.append(np.full(len(nz), j))
        vals.append(c[nz])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(model.N, model.N))


@pytest.mark.parametrize("n,layers", [(8, 3), (16, 6)])
def test_levels_and_vcycle_match_oracle(n, layers):
    import torch
    from mammos_mumag_amd import Model
    from oracle import amg
    g = generate.build(edge=float(n), n=n, layers=layers, grains=4, cone=15.0, seed=3)
    model = Model.from_structured(g, structured=False, precision="mixed")
    K = context_K(model)
    M = amg.AMG(K)
    print(f"[gpu] AMG on the {len(g.xs)}^3-node general mesh: oracle levels {M.sizes()}", flush=True)
    model.amg()
    assert model.amg_levels() == M.sizes()
    rng = np.random.default_rng(4)
    free = ~amg.dirichlet_rows(K)
    for _ in range(2):   # FP32 against FP64: 1.2e-7 measured on the smaller mesh
        r = rng.normal(size=model.N)
        r[~free] = 0.0
        z = model.amg_apply(torch.tensor(r, dtype=torch.float32, device="cuda")).cpu().numpy().astype(np.float64)
        zr = M.apply(r)
        print(f"[gpu] V-cycle relative difference {rel(z, zr):.2e}", flush=True)
        assert rel(z, zr) < 1e-5
    model.close()


def test_mixed_amg_loop_matches_fp64():
    """Three field steps of the cube40 sweep on the general kernels: mixed precision with
    the AMG against plain FP64 (M.h to 1e-6, energies to 1e-6 relative), with a fraction of
    the mixed Jacobi form's PCG iterations."""
    from mammos_mumag_amd import Model, _lib
    g = generate.build(edge=24.0, n=24, layers=8, grains=4, cone=15.0, seed=3)
    p = generate.loop_params("cube40_single")
    out = {}
    for form in ("fp64", "mixed", "mixed+amg"):
        print(f"[gpu] general-path loop, {form}", flush=True)
        model = Model.from_structured(g, structured=False, precision="fp64" if form == "fp64" else "mixed")
        if form == "mixed+amg":
            model.amg()
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
    for form in ("mixed", "mixed+amg"):
        r1 = out[form][0]
        assert np.max(np.abs(r1[:, 1] - r0[:, 1])) < 1e-6, form
        assert np.max(np.abs(r1[:, 0] / r0[:, 0] - 1)) < 1e-6, form
    assert out["mixed+amg"][1]["pcg_iters"] < out["mixed"][1]["pcg_iters"] / 4


def test_amg_error_behaviour():
    """mm_amg's contract: MM_E_STATE on structured or FP64 contexts, when K is already at
    the coarsest size, and for another level bound once set up; MM_E_ARG for a bound
    outside 2..10; 0 restores Jacobi and the same bound re-enables the built hierarchy."""
    from mammos_mumag_amd import Model, _lib
    small = generate.build(edge=2.0, n=2, layers=2, grains=1, cone=0.0, seed=3)   # 343 nodes
    g = generate.build(edge=8.0, n=8, layers=3, grains=4, cone=15.0, seed=3)
    print("[gpu] mm_amg error paths", flush=True)
    for kw in (dict(structured=True, precision="mixed"), dict(structured=False, precision="fp64")):
        model = Model.from_structured(g, **kw)
        with pytest.raises(_lib.MumagError, match=r"\(-4\)"):
            model.amg()
        model.close()
    model = Model.from_structured(small, structured=False, precision="mixed")
    assert model.N <= 500
    with pytest.raises(_lib.MumagError, match=r"\(-4\)"):
        model.amg()
    model.close()
    model = Model.from_structured(g, structured=False, precision="mixed")
    for bad in (1, 11, -1):
        with pytest.raises(_lib.MumagError, match=r"\(-1\)"):
            model.amg(bad)
    assert model.amg_levels() == []
    model.amg(4)
    sizes = model.amg_levels()
    assert len(sizes) >= 2 and sizes[0] == model.N
    with pytest.raises(_lib.MumagError, match=r"\(-4\)"):
        model.amg(5)
    model.amg(0)
    assert model.amg_levels() == []
    model.amg(4)
    assert model.amg_levels() == sizes
    model.close()
