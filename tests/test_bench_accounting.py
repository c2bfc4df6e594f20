"""The bench's whole-step byte accounting (bench.whole_step) on hand-made counts: every
kernel class with its per-launch bytes, for the mixed form with multigrid, the mixed form
with Jacobi and plain FP64 (CPU only)."""
import importlib.util
import pathlib

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_mod", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _res(levels, vcycles, ru=0, corr=0):
    N, Nb, Nv = 1000, 343, 216
    return dict(info=dict(structured=1), n_nodes=N, n_mag_nodes=Nb, n_voxels=Nv, levels=levels,
                stats=dict(pcg_iters=10, pcg_calls=2, tcg_iters=5, fcalls=2, ncg_iters=3),
                timing=dict(vcycles=vcycles, ru_passes=ru, corrections=corr), elapsed=2.0)


def test_mixed_multigrid(bench):
    r = _res([(10, 10, 10), (6, 6, 6), (4, 4, 4)], vcycles=9, ru=4, corr=2)
    w = bench.whole_step(r, "mixed", steps=2)
    cls = {
        "pcg_iterations": 10 * 12 * 4 * 1000,
        "v_cycles": 9 * (108 * 1000 + 72 * (216 + 64)),
        "true_residuals": 4 * 28 * 1000 + 2 * 24 * 1000,
        "truncated_cg": 5 * (34 * 4 * 343 + 2 * 216) + 3 * 3 * 4 * 343,
        "evaluations": 2 * (216 * 343 + 4 * 216),
    }
    total = sum(cls.values())
    assert w["bytes_per_step"] == total // 2
    assert w["by_class_MB_per_step"] == {k: round(v / 2 / 1e6, 1) for k, v in cls.items()}
    assert w["GBs"] == round(total / 2 / 1.0 / 1e9, 1)
    assert w["counts_per_step"] == dict(vcycles=4.5, ru_passes=2.0, corrections=1.0)


def test_mixed_jacobi_and_fp64(bench):
    w = bench.whole_step(_res([], vcycles=-1, ru=3, corr=1), "mixed", steps=1)
    assert "v_cycles" not in w["by_class_MB_per_step"] and w["counts_per_step"]["vcycles"] is None
    f = bench.whole_step(_res([], vcycles=-1), "fp64", steps=1)
    cls = {"pcg_iterations": 10 * 12 * 8 * 1000, "pcg_starts": 2 * 56 * 1000,
           "truncated_cg": 5 * (34 * 8 * 343 + 2 * 216) + 3 * 3 * 8 * 343,
           "evaluations": 2 * (216 * 343 + 4 * 216)}
    assert f["bytes_per_step"] == sum(cls.values())
    r = _res([], vcycles=-1)
    r["info"]["structured"] = 0
    assert bench.whole_step(r, "fp64", steps=1) is None
