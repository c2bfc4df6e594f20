"""The multigrid restatement (oracle/mg.py): its level operator is the Poisson K of the
FP64 oracle on the interior, the V-cycle is symmetric and positive (a valid CG
preconditioner) and cuts the iteration count of CG on K."""
import numpy as np
import pytest

from oracle import mg


@pytest.fixture(scope="module")
def grid():
    from mammos_mumag_amd import generate
    return generate.build(edge=24.0, n=8, layers=3, grains=4, cone=15.0, seed=7)


def test_level_operator_is_poisson_k(grid):
    from oracle.assemble import Model
    O = Model(grid.coords(), grid.tets(), grid.elem_tags(), grid.krn())
    A, face = mg.operator(grid.xs, grid.ys, grid.zs)
    inner = ~face
    K = O.K.tocsr()
    # interior rows/columns agree with the assembled K (faces: K has identity rows/cols)
    d = (A[inner][:, inner] - K[inner][:, inner]).toarray()
    assert np.max(np.abs(d)) <= 1e-12 * np.max(np.abs(K.diagonal()))
    assert np.allclose(A.diagonal()[face], 1.0) and np.allclose(K.diagonal()[face], 1.0)


@pytest.mark.parametrize("n,layers", [(8, 3), (30, 4)])
def test_vcycle_symmetric_positive_and_effective(n, layers):
    """One level (15 nodes per axis) and three (39 -> 23 -> 14)."""
    from mammos_mumag_amd import generate
    grid = generate.build(edge=24.0, n=n, layers=layers, grains=2, cone=15.0, seed=3)
    V = mg.VCycle(grid.xs, grid.ys, grid.zs)
    assert [len(a) for a, _, _ in V.axes] == ([15] if n == 8 else [39, 23, 14])
    A, face = mg.operator(grid.xs, grid.ys, grid.zs)
    rng = np.random.default_rng(0)
    u, v = rng.normal(size=(2, A.shape[0]))
    u[face] = 0
    v[face] = 0
    Mu, Mv = V.apply(u), V.apply(v)
    assert abs(v @ Mu - u @ Mv) <= 1e-12 * abs(u @ Mu)
    assert u @ Mu > 0 and v @ Mv > 0
    # CG on K with the V-cycle against Jacobi: fewer iterations, same answer
    b = rng.normal(size=A.shape[0])
    b[face] = 0

    def pcg(M, tol=1e-10):
        x = np.zeros_like(b)
        r = b.copy()
        z = M(r)
        p = z.copy()
        rz = r @ z
        for it in range(2000):
            if r @ r <= tol * tol * (b @ b):
                return x, it
            q = A @ p
            al = rz / (p @ q)
            x += al * p
            r -= al * q
            z = M(r)
            rz2 = r @ z
            p = z + rz2 / rz * p
            rz = rz2
        return x, 2000

    xj, itj = pcg(lambda r: r / A.diagonal())
    xm, itm = pcg(V.apply)
    assert itm < itj / 2
    assert np.linalg.norm(xm - xj) <= 1e-8 * np.linalg.norm(xj)
