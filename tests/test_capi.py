"""The C ABI library loads and exports what include/mumag_hip.h declares; argument and
mesh validation happen on the host (no GPU needed)."""
import ctypes
import pathlib
import re

import numpy as np
import pytest

from mammos_mumag_amd import _lib

HEADER = pathlib.Path(__file__).resolve().parent.parent / "include" / "mumag_hip.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mm_[A-Za-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_lib.exported_symbols())


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def _desc(coords, tets, mat, Js=(1.0,), keep=None):
    keep = [] if keep is None else keep
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    mat = np.ascontiguousarray(mat, dtype=np.int32)
    nm = len(Js)
    arrs = dict(Js=np.array(Js, dtype=np.float64), A=np.ones(nm), K=np.ones(nm),
                u=np.tile([0.0, 0.0, 1.0], nm))
    keep += [coords, tets, mat, *arrs.values()]
    d = _lib.MeshDesc()
    d.n_nodes, d.n_elems, d.n_mats = len(coords), len(tets), nm
    d.coords = coords.ctypes.data_as(_lib.c_double_p)
    d.tets = tets.ctypes.data_as(_lib.c_int32_p)
    d.elem_mat = mat.ctypes.data_as(_lib.c_int32_p)
    d.mat_Js = arrs["Js"].ctypes.data_as(_lib.c_double_p)
    d.mat_A = arrs["A"].ctypes.data_as(_lib.c_double_p)
    d.mat_K1 = arrs["K"].ctypes.data_as(_lib.c_double_p)
    d.mat_u = arrs["u"].ctypes.data_as(_lib.c_double_p)
    d.length_scale = 1e-9
    return d


def _create(desc):
    lib = _lib.load()
    ctx = ctypes.c_void_p()
    rc = lib.mm_ctx_create(ctypes.byref(desc) if desc is not None else None, None, ctypes.byref(ctx))
    return rc, lib.mm_last_error().decode()


def test_null_arguments_are_rejected():
    rc, msg = _create(None)
    assert rc == _lib.MM_E_ARG and msg
    lib = _lib.load()
    assert lib.mm_ctx_destroy(None) == _lib.MM_E_ARG
    assert lib.mm_total_eg(None, None, 0.0, _lib.D3(0, 0, 1), 1e-10, 1, None, None, None, None, None) == _lib.MM_E_ARG
    assert lib.mm_minimize(None, None, 0.0, _lib.D3(0, 0, 1), None, None, None, None, None, None) == _lib.MM_E_ARG


def test_bad_meshes_are_rejected_before_any_device_work():
    coords = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    keep = []
    rc, msg = _create(_desc(coords, [[0, 1, 2, 7]], [0], keep=keep))
    assert rc == _lib.MM_E_MESH and "out of range" in msg
    rc, msg = _create(_desc(coords, [[0, 1, 2, 2]], [0], keep=keep))
    assert rc == _lib.MM_E_MESH and "degenerate" in msg
    rc, msg = _create(_desc(coords, [[0, 1, 2, 3]], [0], Js=(0.0,), keep=keep))
    assert rc == _lib.MM_E_MESH and "no magnetic element" in msg
    rc, msg = _create(_desc(coords, [[0, 1, 2, 3]], [3], keep=keep))
    assert rc == _lib.MM_E_MESH and "material" in msg


def test_check_raises_with_message():
    with pytest.raises(_lib.MumagError):
        _lib.check(_lib.MM_E_ARG, "probe")
    assert _lib.check(_lib.MM_W_MAXITER) == _lib.MM_W_MAXITER
