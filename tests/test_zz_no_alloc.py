"""mm_dot_noalloc (the context-free, allocation-free dot product) against mm_dot through the
C ABI: bit-identical on two calls in a row, the second with a smaller input (the static
device array keeps the first call's partials; the second must not read them), and the
argument checks.  Runs alone in a fresh process as well as after the other tests.
"""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(lib, a, b, n, device=0):
    out = ctypes.c_double(float("nan"))
    rc = lib.mm_dot_noalloc(device, None if a is None else ctypes.c_void_p(a.data_ptr()),
                            None if b is None else ctypes.c_void_p(b.data_ptr()), n, ctypes.byref(out))
    return rc, out.value


def test_noalloc_dot_matches_mm_dot(cube_arrays, cube_krn):
    import torch
    from mammos_mumag_amd import Model, _lib
    print("[gpu] mm_dot_noalloc vs mm_dot", flush=True)
    lib = _lib.load()
    model = Model(*cube_arrays, cube_krn)
    gen = torch.Generator().manual_seed(11)
    for n in (3_000_001, 12_345):
        a = torch.randn(n, dtype=torch.float64, generator=gen).cuda()
        b = torch.randn(n, dtype=torch.float64, generator=gen).cuda()
        torch.cuda.synchronize()
        rc, got = _noalloc(lib, a, b, n)
        assert rc == 0, lib.mm_last_error()
        want = model.dot(a, b)
        ref = math.fsum((a.cpu().numpy() * b.cpu().numpy()).tolist())
        print(f"[gpu] n={n}: noalloc {got!r} mm_dot {want!r} fsum {ref!r}", flush=True)
        assert got == want                                   # same kernels, same order
        assert abs(got - ref) <= 1e-12 * math.sqrt(n) * float(torch.dot(a.abs(), b.abs()))
    model.close()


def test_noalloc_dot_edge_cases():
    import torch
    from mammos_mumag_amd import _lib
    lib = _lib.load()
    a = torch.ones(1000, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert _noalloc(lib, a, a, 1000) == (0, 1000.0)
    assert _noalloc(lib, None, None, 0) == (0, 0.0)        # empty input: no element is read
    assert _noalloc(lib, a, a, 1) == (0, 1.0)
    assert _noalloc(lib, a, a, -1)[0] == _lib.MM_E_ARG
    assert _noalloc(lib, None, a, 5)[0] == _lib.MM_E_ARG
    assert _noalloc(lib, a, a, 10, device=torch.cuda.device_count())[0] == _lib.MM_E_ARG
    assert _noalloc(lib, a, a, 10, device=-1)[0] == _lib.MM_E_ARG
    out = ctypes.c_double()
    assert lib.mm_dot_noalloc(0, ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(a.data_ptr()), 10, None) == _lib.MM_E_ARG
    assert _noalloc(lib, a, a, 1000) == (0, 1000.0)        # still usable after rejected calls
    del out
