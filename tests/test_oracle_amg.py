"""The smoothed-aggregation restatement (oracle/amg.py): aggregation on a path graph worked
out by hand, aggregation invariants and the smoothed prolongation row by row on a
generated mesh's K, the Galerkin identity, and the V-cycle's symmetry, positivity and cut
of CG iterations against Jacobi."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import amg


def path_laplacian(n):
    """1-D Laplacian on n nodes with Dirichlet (identity) rows at both ends."""
    main = np.full(n, 2.0)
    A = sp.diags([main, -np.ones(n - 1), -np.ones(n - 1)], [0, -1, 1]).tolil()
    for i in (0, n - 1):
        A[i, :] = 0.0
        A[:, i] = 0.0
        A[i, i] = 1.0
    return A.tocsr()


def test_path_graph_aggregates_by_hand():
    """Nodes 1..8 free: pass 1 takes {1, 2} (node 0 is Dirichlet), skips 3 (its neighbour 2
    is taken), takes {3, 4, 5} around 4 and {6, 7, 8} around 7; nothing is left for passes
    2 and 3."""
    A = path_laplacian(10)
    free = ~amg.dirichlet_rows(A)
    assert list(np.flatnonzero(~free)) == [0, 9]
    agg, na = amg.aggregate(amg.strength(A, free, 0.0), free)
    assert na == 3
    assert list(agg) == [-1, 0, 0, 1, 1, 1, 2, 2, 2, -1]
    # pass 2 and 3: on 6 free nodes (1..6), {1, 2} and {3, 4, 5} in pass 1, 6 joins the
    # aggregate of its pass-1 neighbour 5 in pass 2
    A = path_laplacian(8)
    free = ~amg.dirichlet_rows(A)
    agg, na = amg.aggregate(amg.strength(A, free, 0.0), free)
    assert list(agg) == [-1, 0, 0, 1, 1, 1, 1, -1] and na == 2
    # a node with no strong neighbour starts its own aggregate (pass 1, vacuously)
    B = sp.csr_matrix(np.diag([1.0, 2.0, 3.0]))
    B[1, 1] = 2.0
    agg, na = amg.aggregate(amg.strength(B, np.array([True, True, True]), 0.0), np.array([True, True, True]))
    assert list(agg) == [0, 1, 2] and na == 3


@pytest.fixture(scope="module")
def K():
    from mammos_mumag_amd import generate
    from oracle.assemble import Model
    g = generate.build(edge=16.0, n=16, layers=6, grains=4, cone=15.0, seed=3)
    return Model(g.coords(), g.tets(), g.elem_tags(), g.krn()).K.tocsr()


def test_aggregates_and_prolongation(K):
    M = amg.AMG(K)
    lev = M.levels[0]
    A, agg, T, P = lev["A"], lev["agg"], lev["T"], lev["P"]
    free = ~amg.dirichlet_rows(A)
    # every free node in exactly one aggregate, the Dirichlet nodes in none
    assert np.all(agg[free] >= 0) and np.all(agg[~free] == -1)
    na = T.shape[1]
    assert np.array_equal(np.bincount(agg[free], minlength=na), np.asarray(T.sum(axis=0)).ravel())
    # each aggregate is connected through stored couplings between its own nodes
    S = amg.strength(A, free, 0.0)
    for a in np.random.default_rng(0).choice(na, 40, replace=False):
        nodes = set(np.flatnonzero(agg == a))
        seen, todo = set(), [min(nodes)]
        while todo:
            i = todo.pop()
            if i in seen:
                continue
            seen.add(i)
            todo += [j for j in S.indices[S.indptr[i]:S.indptr[i + 1]] if j in nodes]
        assert seen == nodes
    # P = (I - w D^-1 A) T row by row (dense), w = 4 / (3 rho), rho the Gershgorin bound
    rho = np.max(np.abs(A).sum(axis=1).A1 / A.diagonal())
    assert lev["rho"] == pytest.approx(rho, rel=1e-15)
    w = 4.0 / (3.0 * rho)
    Td = T.toarray()
    for i in np.random.default_rng(1).choice(A.shape[0], 60, replace=False):
        row = Td[i] - w / A[i, i] * (A[i].toarray().ravel() @ Td)
        assert np.allclose(P[i].toarray().ravel(), row, rtol=0, atol=1e-15)
    # Dirichlet rows of P are zero
    assert P[np.flatnonzero(~free)].nnz == 0


def test_galerkin_vcycle_and_cg(K):
    M = amg.AMG(K)
    assert len(M.levels) >= 3 and M.sizes()[-1] <= 500
    for lev, nxt in zip(M.levels[:-1], M.levels[1:]):
        Ac = (lev["P"].T @ lev["A"] @ lev["P"]).toarray()
        assert np.allclose(nxt["A"].toarray(), Ac, rtol=0, atol=1e-13 * np.abs(Ac).max())
        assert abs(nxt["A"] - nxt["A"].T).max() <= 1e-13 * abs(nxt["A"]).max()
    free = ~amg.dirichlet_rows(K)
    rng = np.random.default_rng(2)
    u, v, b = rng.normal(size=(3, K.shape[0]))
    for x in (u, v, b):
        x[~free] = 0.0
    Mu, Mv = M.apply(u), M.apply(v)
    assert abs(v @ Mu - u @ Mv) <= 1e-12 * abs(u @ Mu)
    assert u @ Mu > 0 and v @ Mv > 0
    assert np.all(Mu[~free] == 0.0)

    def pcg(Mf, tol=1e-10):
        x = np.zeros_like(b)
        r = b.copy()
        z = Mf(r)
        p = z.copy()
        rz = r @ z
        for it in range(3000):
            if r @ r <= tol * tol * (b @ b):
                return x, it
            q = K @ p
            al = rz / (p @ q)
            x += al * p
            r -= al * q
            z = Mf(r)
            rz2 = r @ z
            p = z + rz2 / rz * p
            rz = rz2
        return x, 3000

    dinv = 1.0 / K.diagonal()
    xj, itj = pcg(lambda r: dinv * r)
    xm, itm = pcg(M.apply)
    assert itm < itj / 5
    assert np.linalg.norm(xm - xj) <= 1e-8 * np.linalg.norm(xj)


def test_strength_threshold_follows_anisotropy():
    """On a 2D Laplacian with x couplings -1 and y couplings -0.01 (s_ij = 0.01 / 2.02 below
    the threshold 0.02 across rows, 1 / 2.02 along them) every aggregate stays within one
    grid row (semi-coarsening along the strong direction); with theta = 0 every coupling
    is strong and the aggregates span rows."""
    nx, ny = 12, 6
    idx = np.arange(nx * ny).reshape(ny, nx)
    rows, cols, vals = [], [], []
    for j in range(ny):
        for i in range(nx):
            p = idx[j, i]
            rows.append(p), cols.append(p), vals.append(2.02)
            for di, dj, w in ((1, 0, -1.0), (-1, 0, -1.0), (0, 1, -0.01), (0, -1, -0.01)):
                if 0 <= i + di < nx and 0 <= j + dj < ny:
                    rows.append(p), cols.append(idx[j + dj, i + di]), vals.append(w)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(nx * ny, nx * ny))
    free = np.ones(nx * ny, dtype=bool)
    row_of = np.repeat(np.arange(ny), nx)
    agg, na = amg.aggregate(amg.strength(A, free, 0.02), free)
    assert np.all(agg >= 0)
    for a in range(na):
        assert len(set(row_of[agg == a])) == 1
    agg0, na0 = amg.aggregate(amg.strength(A, free, 0.0), free)
    assert any(len(set(row_of[agg0 == a])) > 1 for a in range(na0))
    assert na0 < na
