"""The oracle through a magnetisation reversal, pinned by the reference's own record of
one: the quickstart loop of examples/quickstart.ipynb (hysteresis.run on CUBE_20_nm =
tests/data/cube.fly, Ms 1.28 MA/m, A 7.7 pJ/m, K1 4.3 MJ/m^3, +10 T to -10 T in 20
steps; 17 records, M reverses between -5 T and -6 T).  The notebook prints M and the
energy density to ~15 digits; the oracle must reproduce them to 1e-12 relative, with
the same snapshot counters.

Also: the oracle reproduces its own recorded branch cases (tests/golden/oracle_branches.json,
written by tests/golden/make_switching.py), which the GPU tests compare against; they
exercise the line-search halvings and both descent restarts of the NCG.
"""
import json

import numpy as np
import pytest

from golden.make_switching import COUNTERS, branch_cases, quickstart_inputs, run_oracle_case


@pytest.fixture(scope="module")
def qs(golden):
    return json.loads((golden / "quickstart_loop.json").read_text())


@pytest.fixture(scope="module")
def branches(golden):
    return json.loads((golden / "oracle_branches.json").read_text())


def test_notebook_fixture_shape(qs):
    assert len(qs["H"]) == len(qs["M"]) == len(qs["energy_density"]) == 17
    assert qs["configuration_type"] == [1] * 16 + [2]
    assert qs["H_unit"] == "A / m" and qs["M_unit"] == "A / m" and qs["energy_density_unit"] == "J / m3"
    # reversal between the last two records: M flips sign, |M| stays near Ms
    M = np.array(qs["M"])
    assert np.all(M[:16] > 0) and M[16] < 0
    assert np.all(np.abs(M) / 1.28e6 > 0.99)


def test_reconstructed_field_values(qs):
    """hstart / hstep reconstructed from the notebook's call reproduce its printed H."""
    q = quickstart_inputs()
    assert q == qs["inputs"]
    h = q["hstart"]
    H = []
    for _ in range(17):
        H.append(h / qs["mu0_astropy"])
        h = h + q["hstep"]
    assert np.max(np.abs(np.array(H) - np.array(qs["H"]))) < 1e-6   # printed with 8 decimals


def test_oracle_reverses_like_the_notebook(golden, qs, cube_arrays):
    from oracle.io import mu0_output
    r = run_oracle_case("quickstart", *cube_arrays)
    rec = np.array(r["records"])
    mu0a = qs["mu0_astropy"]
    assert rec.shape == (17, 4)
    assert np.array_equal(rec[:, 0].astype(int), qs["configuration_type"])
    assert np.max(np.abs(rec[:, 1] / mu0a - qs["H"])) < 1e-6
    assert np.max(np.abs(rec[:, 2] / mu0a / qs["M"] - 1)) < 1e-12
    assert np.max(np.abs(rec[:, 3] / mu0_output() / qs["energy_density"] - 1)) < 1e-12


@pytest.mark.parametrize("name", ["vortex_hmag0", "vortex_hmag1", "tilt1"])
def test_oracle_branch_cases_reproduce(name, branches, cube_arrays):
    r = run_oracle_case(name, *cube_arrays)
    ref = branches[name]
    assert r["counters"] == ref["counters"]
    assert np.max(np.abs(np.array(r["records"]) - np.array(ref["records"]))) < 1e-12


def test_branch_cases_take_the_branches(branches):
    """The cases cover what the reference's test loop never reaches."""
    last = {k: v["counters"][-1] for k, v in branches.items()}
    assert last["vortex_hmag0"]["backtracks"] == 8 and last["vortex_hmag1"]["backtracks"] == 1
    assert last["tilt1"]["restarts_g"] == 16 and last["tilt1"]["restarts_z"] == 17
    assert last["quickstart"]["backtracks"] == 0
    assert set(COUNTERS) <= set(last["quickstart"])
    recs = np.array(branches["tilt1"]["records"])
    assert recs[-2, 2] > 0 > recs[-1, 2]         # reversal at the last step (-4 T)
    assert len(branch_cases()) == len(branches)
