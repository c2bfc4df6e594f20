"""Test configuration: markers, import paths and shared fixtures."""
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "mammos-mumag_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


@pytest.fixture(scope="session")
def cube_arrays():
    """The reference test mesh (tests/data/cube.fly): coords, tets, element tags."""
    z = np.load(GOLDEN / "cube_mesh.npz")
    return z["coords"], z["tets"], z["elem_tags"]


@pytest.fixture(scope="session")
def cube_krn(cube_arrays):
    from oracle.io import read_krn
    tags = sorted(set(int(t) for t in np.unique(cube_arrays[2])))
    return read_krn(str(GOLDEN / "cube.krn"), tags)


@pytest.fixture(scope="session")
def cube_params():
    from oracle.io import read_p2
    return read_p2(str(GOLDEN / "cube.p2"))


@pytest.fixture(scope="session")
def oracle_cube(cube_arrays, cube_krn):
    from oracle.assemble import Model
    return Model(*cube_arrays, cube_krn)


@pytest.fixture(scope="session")
def cube_dir(tmp_path_factory, cube_arrays):
    """A directory holding cube.fly / cube.krn / cube.p2 as the reference's tests use them."""
    import shutil
    from mammos_mumag_amd.fileio import write_fly
    d = tmp_path_factory.mktemp("cube_inputs")
    write_fly(d / "cube.fly", *cube_arrays)
    shutil.copyfile(GOLDEN / "cube.krn", d / "cube.krn")
    shutil.copyfile(GOLDEN / "cube.p2", d / "cube.p2")
    return d


def read_csv_values(path):
    import csv
    with open(path) as f:
        rows = list(csv.reader(f))[2:]
    return np.array([float(r[1]) for r in rows if r])
