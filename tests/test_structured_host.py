"""Generator invariants and the host-side validation of structured (Kuhn-grid) contexts.

The validation in mm_ctx_create runs before any device call, so a wrong description is
rejected with MM_E_MESH here on the CPU; a correct one gets past validation (and then
either builds a context on a GPU box or stops at the missing device).
"""
import ctypes

import numpy as np
import pytest

from mammos_mumag_amd import _lib, generate
from mammos_mumag_amd.model import boundary_setup


@pytest.fixture(scope="module")
def small():
    return generate.build(edge=24.0, n=8, layers=3, grains=4, cone=15.0, seed=7)


def test_generator_geometry(small):
    g = small
    X, T, tags = g.coords(), g.tets(), g.elem_tags()
    nx = len(g.xs)
    assert len(X) == nx ** 3 and len(T) == 6 * (nx - 1) ** 3
    X0 = X[T[:, 0]]
    det = np.einsum("ij,ij->i", X[T[:, 1]] - X0, np.cross(X[T[:, 2]] - X0, X[T[:, 3]] - X0))
    assert det.min() > 0  # positive orientation
    vol = np.abs(det) / 6
    assert np.isclose(vol.sum(), (3 * 24.0) ** 3, rtol=1e-12)
    # the air layers span exactly one magnet edge on each side
    assert g.xs[0] == -36.0 and g.xs[-1] == 36.0 and g.xs[3] == -12.0 and g.xs[11] == 12.0
    assert np.allclose(np.diff(g.xs[3:12]), 3.0)
    # air box branch of the reference's set-up: u = 0 on the six faces
    b = boundary_setup(X, T, tags, sorted(set(tags.tolist())))
    assert b["airbox"]
    face = (np.abs(X) == 36.0).any(axis=1)
    assert np.array_equal(b["dirichlet"], face)
    # grain boundary phase: every grain-grain face contact goes through it
    vt = g.voxel_tag[3:11, 3:11, 3:11]
    G = g.meta["grains"]
    for ax in range(3):
        a = np.moveaxis(vt, ax, 0)
        pairs = np.stack([a[:-1], a[1:]], -1).reshape(-1, 2)
        bad = (pairs[:, 0] != pairs[:, 1]) & (pairs[:, 0] <= G) & (pairs[:, 1] <= G)
        assert not bad.any()
    # .krn lines follow the sorted tags; air is last and non-magnetic
    tags_sorted = sorted(g.tags_props)
    assert g.tags_props[tags_sorted[-1]]["Js"] == 0.0
    assert all(g.tags_props[t]["Js"] > 0 for t in tags_sorted[:-1])


def _desc(g, keep, **override):
    X, T, tags = g.coords(), g.tets(), g.elem_tags()
    tags_sorted = sorted(g.tags_props)
    mat = np.searchsorted(tags_sorted, tags).astype(np.int32)
    props = [g.tags_props[t] for t in tags_sorted]
    Js = np.array([p["Js"] for p in props])
    A = np.array([p["A"] for p in props])
    K1 = np.array([p["K1"] for p in props])
    u = np.array([[np.sin(p["theta"]) * np.cos(p["phi"]), np.sin(p["theta"]) * np.sin(p["phi"]),
                   np.cos(p["theta"])] for p in props]).ravel()
    dirichlet = (np.abs(X) == np.abs(X).max()).any(axis=1).astype(np.uint8)
    arrs = dict(X=X, T=T, mat=mat, Js=Js, A=A, K1=K1, u=u, dirichlet=dirichlet,
                xs=g.xs.copy(), ys=g.ys.copy(), zs=g.zs.copy())
    arrs.update(override)
    arrs = {k: np.ascontiguousarray(v) for k, v in arrs.items()}
    arrs["T"] = np.ascontiguousarray(arrs["T"], dtype=np.int32)
    keep.append(arrs)
    d = _lib.MeshDesc()
    d.n_nodes, d.n_elems, d.n_mats = len(arrs["X"]), len(arrs["T"]), len(Js)
    d.coords = arrs["X"].ctypes.data_as(_lib.c_double_p)
    d.tets = arrs["T"].ctypes.data_as(_lib.c_int32_p)
    d.elem_mat = arrs["mat"].ctypes.data_as(_lib.c_int32_p)
    d.mat_Js = arrs["Js"].ctypes.data_as(_lib.c_double_p)
    d.mat_A = arrs["A"].ctypes.data_as(_lib.c_double_p)
    d.mat_K1 = arrs["K1"].ctypes.data_as(_lib.c_double_p)
    d.mat_u = arrs["u"].ctypes.data_as(_lib.c_double_p)
    d.dirichlet = arrs["dirichlet"].ctypes.data_as(_lib.c_uint8_p)
    d.length_scale = 1e-9
    nv = len(g.xs) - 1
    d.grid_n[0] = d.grid_n[1] = d.grid_n[2] = nv
    d.grid_x = arrs["xs"].ctypes.data_as(_lib.c_double_p)
    d.grid_y = arrs["ys"].ctypes.data_as(_lib.c_double_p)
    d.grid_z = arrs["zs"].ctypes.data_as(_lib.c_double_p)
    m0, mn = g.magnet_box[0]
    d.grid_mag0[0] = d.grid_mag0[1] = d.grid_mag0[2] = override.get("mag0", m0)
    d.grid_magn = override.get("magn", mn)
    return d


def _create(d):
    lib = _lib.load()
    opts = _lib.Options()
    opts.use_structured = 1
    ctx = ctypes.c_void_p()
    rc = lib.mm_ctx_create(ctypes.byref(d), ctypes.byref(opts), ctypes.byref(ctx))
    msg = lib.mm_last_error().decode()
    if rc == 0:
        lib.mm_ctx_destroy(ctx)
    return rc, msg


def test_valid_description_passes_validation(small):
    keep = []
    rc, msg = _create(_desc(small, keep))
    assert rc == 0 or (rc == _lib.MM_E_HIP and "hipSetDevice" in msg), msg


def test_invalid_descriptions_are_rejected(small):
    keep = []
    g = small
    T = g.tets().copy()
    T[10, [1, 2]] = T[10, [2, 1]]  # not the generator's Kuhn order
    rc, msg = _create(_desc(g, keep, T=T))
    assert rc == _lib.MM_E_MESH and "Kuhn" in msg
    rc, msg = _create(_desc(g, keep, mag0=2))  # wrong magnet block
    assert rc == _lib.MM_E_MESH and "magnet block" in msg
    X = g.coords().copy()
    X[5, 0] += 1e-3
    rc, msg = _create(_desc(g, keep, X=X))
    assert rc == _lib.MM_E_MESH and "grid position" in msg
    d = np.zeros(len(g.coords()), dtype=np.uint8)
    rc, msg = _create(_desc(g, keep, dirichlet=d))
    assert rc == _lib.MM_E_MESH and "Dirichlet" in msg
