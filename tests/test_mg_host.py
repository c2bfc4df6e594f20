"""Host-side guards of the multigrid set-up (no GPU):

* ``mg_restrict_span_fits`` (mammos-mumag_amd/csrc/mg.h), which mm_multigrid applies to
  every level's x restriction lists, accepts lists whose coarse intervals cover at most two
  fine ones (and a few of three, within the staged row's slack) and rejects lists that
  reach past the restriction kernel's staged row: compiled from the header with hipcc
  (host code only) and run here (tests/host/mg_span_check.cpp).
* The restriction pairs of the spacing-aware coarsening of the benchmark grids, built from
  the oracle's interpolation (oracle/mg.py interp, transposed), meet the kernel's own
  condition: for each 64 consecutive coarse nodes every fine index with a nonzero weight
  lies in [first fine index of the first one, + 132).
"""
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

from oracle import mg

ROOT = pathlib.Path(__file__).resolve().parent.parent
CSRC = ROOT / "mammos-mumag_amd" / "csrc"


def test_span_guard_cases(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not pathlib.Path(hipcc).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path / "mg_span_check"
    subprocess.run([hipcc, "-std=c++17", "-O1", "-x", "hip", "--offload-arch=gfx950", f"-I{CSRC}", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "host" / "mg_span_check.cpp"), "-o", str(exe)], check=True,
                   capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    rows = [line.split() for line in out.stdout.splitlines()]
    assert len(rows) == 5
    for name, want, got in rows:
        assert want == got, name
    assert out.returncode == 0
    assert {r[0]: r[1] for r in rows}["seven_three_fine_intervals"] == "0"


def _pairs(xf, keep):
    """Per coarse index the (fine index, weight) pairs of the transposed interpolation."""
    Pt = mg.interp(xf, keep).T.tocsr()
    out = []
    for C in range(Pt.shape[0]):
        row = Pt.getrow(C)
        out.append([(int(i), float(w)) for i, w in zip(row.indices, row.data) if w != 0.0])
    return out


@pytest.mark.parametrize("which", ["voronoi64_250nm", "graded39"])
def test_benchmark_restriction_lists_fit_the_staged_row(which):
    from mammos_mumag_amd import generate
    g = (generate.build_config("voronoi64_250nm", seed=1) if which == "voronoi64_250nm"
         else generate.build(edge=39.0, n=39, layers=13, grains=2, cone=15.0, seed=3))
    axes = [np.asarray(g.xs, float), np.asarray(g.ys, float), np.asarray(g.zs, float)]
    levels = 0
    while max(len(a) for a in axes) > mg.COARSEST:
        hmin = min(np.diff(a).min() for a in axes)
        keeps = [mg.coarsen(a, hmin) for a in axes]
        if all(len(k) == len(a) for k, a in zip(keeps, axes)):
            break
        pairs = _pairs(axes[0], keeps[0])
        assert max(len(p) for p in pairs) <= 3                     # the kernel's 3 slots
        for i0 in range(0, len(pairs), 64):
            base = pairs[i0][0][0]
            used = [i for p in pairs[i0:i0 + 64] for i, _ in p]
            assert min(used) >= base and max(used) - base < 132, (levels, i0)
        axes = [a[k] for a, k in zip(axes, keeps)]
        levels += 1
    assert levels >= 2
