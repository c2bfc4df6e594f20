"""Initial magnetisation states (scripts/magnetization.py:9-71 getM) by their geometry,
on the reference's test mesh: unit length on the magnet, zero off it, and the defining
property of each state — the vortex circulates about +x with its core along +x, the
flower's tilt x z / (10 s) is radial and odd in z, the twisted state adds an azimuthal
turn 4 z / s (odd in z) to the flower.  The golden generator's own vortex restatement
(tests/golden/make_switching.py, used for the oracle's branch cases) is the same state.
"""
import numpy as np
import pytest

from mammos_mumag_amd.model import initial_state


@pytest.fixture(scope="module")
def mesh(oracle_cube, cube_arrays):
    coords = cube_arrays[0]
    return coords, oracle_cube.meas > 0


def _unit_on_magnet(m, magnet):
    m = m.reshape(-1, 3)
    assert np.allclose(np.linalg.norm(m[magnet], axis=1), 1.0, atol=1e-14)
    assert np.all(m[~magnet] == 0)
    return m


def test_uniform(mesh):
    coords, magnet = mesh
    m = _unit_on_magnet(initial_state(coords, magnet, [0.0, 3.0, 4.0], 0), magnet)
    assert np.allclose(m[magnet], [0.0, 0.6, 0.8], atol=1e-15)


def test_vortex(mesh, oracle_cube):
    coords, magnet = mesh
    m = _unit_on_magnet(initial_state(coords, magnet, [0, 0, 1], 2), magnet)
    x = coords[magnet]
    mm = m[magnet]
    r = np.hypot(x[:, 1], x[:, 2])
    rc = 0.14 * np.hypot(coords[:, 1], coords[:, 2]).max()
    # the in-plane part is tangential and turns counter-clockwise about +x
    assert np.max(np.abs(mm[:, 1] * x[:, 1] + mm[:, 2] * x[:, 2])) < 1e-12 * np.max(r)
    off = r > 1e-9
    assert np.all(mm[off, 2] * x[off, 1] - mm[off, 1] * x[off, 2] >= 0)
    # core along +x, in-plane far from the axis (exp(-2 r / rc) <= e^-8 beyond 4 rc)
    assert np.all(mm[:, 0] > 0)
    assert np.all(mm[r > 4 * rc, 0] < 1e-3)
    near = r < 0.05 * rc
    if near.any():
        assert np.all(mm[near, 0] > 0.85)
    from golden.make_switching import vortex_state
    assert np.allclose(vortex_state(oracle_cube), m.ravel(), atol=1e-15)


def test_flower(mesh):
    coords, magnet = mesh
    m = _unit_on_magnet(initial_state(coords, magnet, [0, 0, 1], 1), magnet)
    x, mm = coords[magnet], m[magnet]
    s = coords.max()
    assert np.all(mm[:, 2] > 0)
    # radial tilt: (mx, my) parallel to (x, y) and of size |z| rho / (10 s) relative to mz
    assert np.max(np.abs(mm[:, 0] * x[:, 1] - mm[:, 1] * x[:, 0])) < 1e-12 * s
    rho = np.hypot(x[:, 0], x[:, 1])
    tilt = np.hypot(mm[:, 0], mm[:, 1]) / mm[:, 2]
    assert np.allclose(tilt, np.abs(x[:, 2]) * rho / (10 * s), atol=1e-13)
    # outward above the mid-plane, inward below
    assert np.all((mm[:, 0] * x[:, 0] + mm[:, 1] * x[:, 1]) * x[:, 2] >= -1e-15)


def test_twisted(mesh):
    coords, magnet = mesh
    m = _unit_on_magnet(initial_state(coords, magnet, [0, 0, 1], 3), magnet)
    f = initial_state(coords, magnet, [0, 0, 1], 1).reshape(-1, 3)
    x, mm, ff = coords[magnet], m[magnet], f[magnet]
    s = coords.max()
    rho = np.hypot(x[:, 0], x[:, 1])
    ok = rho > 1e-9
    ephi = np.stack([-x[:, 1], x[:, 0]], axis=1)[ok] / rho[ok, None]
    erho = np.stack([x[:, 0], x[:, 1]], axis=1)[ok] / rho[ok, None]
    # azimuthal turn over m_z is 4 z / s (odd in z); the radial part is the flower's
    turn = np.sum(mm[ok, :2] * ephi, axis=1) / mm[ok, 2]
    assert np.allclose(turn, 4 * x[ok, 2] / s, atol=1e-12)
    radial = np.sum(mm[ok, :2] * erho, axis=1) / mm[ok, 2]
    fradial = np.sum(ff[ok, :2] * erho, axis=1) / ff[ok, 2]
    assert np.allclose(radial, fradial, atol=1e-12)


def test_random_is_rejected(mesh):
    coords, magnet = mesh
    with pytest.raises(ValueError):
        initial_state(coords, magnet, [0, 0, 1], 4)
