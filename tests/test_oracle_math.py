"""Closed forms, invariants and brute force that pin the oracle where fixtures do not."""
import numpy as np
import scipy.sparse.linalg as spla

from oracle.assemble import KEAST_A, KEAST_B, Model, element_geometry, t3d_tensor
from oracle.minimizers import alpha_init, compute_h, normalize_vectors
from oracle.solvers import jax_cg, truncated_cg


def _unit_tet():
    coords = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    return coords, np.array([[0, 1, 2, 3]], dtype=np.int32)


def test_unit_tet_stiffness():
    """P1 stiffness of the unit tetrahedron: (1/6)[[3,-1,-1,-1],[-1,1,0,0],...]."""
    coords, tets = _unit_tet()
    vol, G = element_geometry(coords, tets)
    assert np.isclose(vol[0], 1 / 6)
    S = vol[0] * G[0] @ G[0].T
    ref = np.array([[3, -1, -1, -1], [-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]]) / 6
    assert np.allclose(S, ref, atol=1e-15)


def test_keast_rule_exact_degree2():
    """The 4-point rule integrates x^2 over the unit tet exactly (1/60)."""
    coords, _ = _unit_tet()
    lam = np.full((4, 4), KEAST_B)
    np.fill_diagonal(lam, KEAST_A)
    xq = lam @ coords
    assert np.isclose(np.sum(xq[:, 0] ** 2) / 4 / 6, 1 / 60, rtol=1e-14)
    assert np.isclose(np.sum(xq[:, 0] * xq[:, 1]) / 4 / 6, 1 / 120, rtol=1e-14)


def test_t3d_symmetry_and_radial():
    """Imhoff tensor: symmetric; radial eigenvalue a2/x^2 (the f-term vanishes radially)."""
    rng = np.random.default_rng(0)
    xq = rng.normal(size=(10, 3))
    xq *= (20 + 10 * rng.random(10))[:, None] / np.linalg.norm(xq, axis=1)[:, None]
    k = t3d_tensor(xq, 18.0, 32.4)
    assert np.allclose(k, np.transpose(k, (0, 2, 1)))
    n = xq / np.linalg.norm(xq, axis=1)[:, None]
    kn = np.einsum("eab,eb->ea", k, n)
    x = np.linalg.norm(xq, axis=1)
    assert np.allclose(kn, (18.0 * (32.4 - 18.0) / x ** 2)[:, None] * n)


def test_operator_invariants(oracle_cube):
    M = oracle_cube
    # symmetry
    assert abs(M.K - M.K.T).max() < 1e-14 * abs(M.K).max()
    assert abs(M.C - M.C.T).max() < 1e-14 * abs(M.C).max()
    # G = D^T / V
    for a in range(3):
        assert abs(M.Gxyz[a] - M.Dxyz[a].T / M.V).max() == 0
    # K 1 = 0 on rows with no Dirichlet neighbour
    K = M.K
    touching = np.zeros(M.N, dtype=bool)
    touching[M.tets[M.dirichlet[M.tets].any(axis=1)].ravel()] = True
    r = K @ np.ones(M.N)
    assert np.max(np.abs(r[~touching & ~M.dirichlet])) < 1e-12 * np.max(np.abs(K.diagonal()))
    # meas sums to Js V (single magnetic material)
    assert np.isclose(M.meas.sum(), 1.61 * M.V, rtol=1e-14)
    # divergence of a uniform field is a pure surface term: sum_i b_i = 0
    m = M.uniform_m([0.3, -0.2, 0.9])
    b = M.Dxyz[0] @ m[0::3] + M.Dxyz[1] @ m[1::3] + M.Dxyz[2] @ m[2::3]
    assert abs(b.sum()) < 1e-12 * np.abs(b).max()


def test_jax_cg_matches_direct_solve(oracle_cube):
    M = oracle_cube
    rng = np.random.default_rng(1)
    b = rng.normal(size=M.N)
    x, it = jax_cg(M.K, M.diagK, b, np.zeros(M.N), tol=1e-14, maxiter=20000)
    xd = spla.spsolve(M.K.tocsc(), b)
    assert np.max(np.abs(x - xd)) < 1e-9 * np.max(np.abs(xd))
    # converged start: zero iterations, x0 returned untouched
    x2, it2 = jax_cg(M.K, M.diagK, b, xd.copy(), tol=1e-6)
    assert it2 == 0 and np.array_equal(x2, xd)
    # maxiter caps the loop
    _, it3 = jax_cg(M.K, M.diagK, b, np.zeros(M.N), tol=1e-14, maxiter=7)
    assert it3 == 7


def _dense_pcg_reference(B, g, D, n_iter):
    """Textbook PCG for B p = g from p = 0 (the truncated CG without its exits)."""
    p = np.zeros_like(g)
    r = -g.copy()
    z = r / D
    d = -z
    for _ in range(n_iter):
        Bd = B @ d
        a = (r @ z) / (d @ Bd)
        p = p + a * d
        rn = r + a * Bd
        zn = rn / D
        d = -zn + (rn @ zn) / (r @ z) * d
        r, z = rn, zn
    return p


def test_truncated_cg_brute_force():
    rng = np.random.default_rng(2)
    n = 6
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    B = Q @ np.diag(np.logspace(0, 4, n)) @ Q.T
    D = np.diag(B).copy()
    g = 3.0 * rng.normal(size=n)  # |g| > 1: eps = 0.5; r.z >> the 1e-16 guard of beta
    p, it = truncated_cg(lambda v: B @ v, -g, D, max_iter=3)
    assert it == 3
    assert np.allclose(p, _dense_pcg_reference(B, g, D, 3), rtol=1e-12, atol=0)
    # run to convergence: p solves B p = g to |B p - g| <= min(0.5, |g|^1.5)
    p, it = truncated_cg(lambda v: B @ v, -g, D, max_iter=50)
    assert it < 50
    assert np.linalg.norm(B @ p - g) <= 0.5
    # negative curvature at the first step: p stays 0 -> -grad_fk = g is returned
    Bn = -np.eye(n)
    p, it = truncated_cg(lambda v: Bn @ v, -g, np.ones(n), max_iter=5)
    assert it == 1 and np.array_equal(p, g)


def test_line_search_scalars():
    x = np.array([0.0, 0.0, 1.0, 1.0, 0.0, 0.0])
    d = np.array([0.1, 0.0, 0.0, 0.0, -0.2, 0.0])
    eps = np.finfo(float).eps
    t1 = 2 * np.sqrt(eps) * (1 + np.sqrt(2)) / np.sqrt(0.05)
    t2 = np.sqrt(eps) / 0.2
    assert compute_h(x, d) == min(t1, t2)
    g = np.array([-1.0, 0, 0, 0, 1, 0])
    g1 = 0.5 * g
    # first iteration: f0 = max float -> term2 = inf -> secant term1
    assert alpha_init(np.finfo(float).max, 1.0, g, d, g1, 1e-3) == -1e-3 * ((g @ d) / ((g1 @ d) - (g @ d)))
    # negative curvature -> 0.1 |term1|
    assert np.isclose(alpha_init(np.finfo(float).max, 1.0, g, d, 2 * g, 1e-3), 0.1 * 1e-3)
    v = normalize_vectors(np.array([3.0, 0, 4, 0, 0, 0]))
    assert np.allclose(v, [0.6, 0, 0.8, 0, 0, 0])


def test_airbox_detection_on_box():
    """A box mesh centred at the origin takes the airbox branch (8 corner nodes at Rinf)."""
    from oracle.io import MU0_SCRIPT  # noqa: F401
    n = 3
    xs = np.linspace(-1.0, 1.0, n + 1)
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    idx = np.arange((n + 1) ** 3).reshape(n + 1, n + 1, n + 1)
    tets = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                c = [idx[i + a, j + b, k + cc] for a in (0, 1) for b in (0, 1) for cc in (0, 1)]
                v000, v100, v010, v110 = c[0], c[4], c[2], c[6]
                v001, v101, v011, v111 = c[1], c[5], c[3], c[7]
                for path in ([v100, v110], [v100, v101], [v010, v110], [v010, v011],
                             [v001, v101], [v001, v011]):
                    tets.append([v000, path[0], path[1], v111])
    tets = np.array(tets, dtype=np.int32)
    vol, _ = element_geometry(coords, tets)
    assert np.isclose(vol.sum(), 8.0)
    tags = np.where(np.abs(coords[tets].mean(1)).max(1) < 0.34, 1, 2).astype(np.int32)
    krn = {1: dict(theta=0.0, phi=0.0, K1=1e6, Js=1.0, A=1e-11), 2: dict(theta=0, phi=0, K1=0, Js=0, A=0)}
    M = Model(coords, tets, tags, krn)
    assert M.airbox
    assert M.dirichlet.sum() == (n + 1) ** 3 - (n - 1) ** 3
