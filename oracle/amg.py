"""Smoothed-aggregation algebraic multigrid for the general (unstructured) meshes' inner
Poisson solve: CPU restatement of mammos-mumag_amd/csrc/amg.cpp (set-up) and amg.hip
(V-cycle) in FP64, for the GPU checks.  Test infrastructure only.

No counterpart in the reference (its potential solve is Jacobi-preconditioned,
scripts/solvers.py:222-231).  Set-up, per level (matrix A, diagonal D):

* Dirichlet rows of the finest level (no non-zero off-diagonal entry) take no part: they join no aggregate and their rows of P are zero.
* Strength: j is a strong neighbour of i (both free, j != i) when
  s_ij = |a_ij| / sqrt(a_ii a_jj) > max(theta, 1e-10) (entries that cancel to rounding
  level in a Galerkin product never count, whichever way they rounded); theta = 0.02 by
  default.  PCG iterations to 1e-10 from zero, theta 0 / 0.02 / 0.05: 29 / 19 / 17 on a
  graded air-box mesh of 117,649 nodes (operator complexity 1.57 / 1.75 / 2.22), 113 / 29 /
  20 on the 103,823-node spherical-shell mesh (1.20 / 1.66 / 2.07); theta = 0.1 grows the
  coarse operators 10-14x.
* Aggregation in three passes over the nodes in index order: (1) a node whose strong
  neighbours all are still unaggregated starts an aggregate with them; (2) a node left
  over joins the pass-1 aggregate of its strongest aggregated neighbour (largest s_ij;
  values within 1e-9 relative count as ties, which go to the lower index, so that the
  rounding of a Galerkin product cannot decide them); (3) a node still left over starts an
  aggregate with its strong neighbours that are still unaggregated.
* Tentative prolongation T (T_ia = 1 when node i is in aggregate a), smoothed:
  P = (I - w D^-1 A) T with w = 4 / (3 rho), rho the Gershgorin bound of D^-1 A
  (max_i sum_j |a_ij| / a_ii, which bounds its spectrum from above).
* A_c = P^T A P (exact zeros dropped).  Down to at most `coarse` unknowns (or until the
  coarsening stalls, fewer than 20% removed); the last level's matrix is inverted.
  Jacobi takes 313 and 550 iterations on the two meshes above.

V-cycle: on each level but the last, `deg` (level 0; `deg_coarse` on the coarser levels,
1 by default: one PCG iteration more per solve than 2 there, for ~19% less work per
V-cycle) Chebyshev-Richardson steps
x += w_k D^-1 (b - A x), w_k the reciprocal Chebyshev nodes of [rho / alpha, rho] (largest
node first; the first step from x = 0), the residual restricted by P^T, the coarse
correction P x_c added, the same steps in reverse order; the last level x = A^-1 b.  The
steps are polynomials in D^-1 A, so the V-cycle is symmetric (and positive, since every
step's factor has norm below one on the spectrum).

Pinned by: exact aggregation/prolongation properties on small matrices (every free node in
exactly one aggregate, aggregates connected through strong links, T's column sums, P
reproducing what (I - w D^-1 A) T gives row by row), the Galerkin identity, the V-cycle's
symmetry and positivity, and its cut of CG iterations against Jacobi
(tests/test_oracle_amg.py).
"""
import numpy as np
import scipy.sparse as sp


def dirichlet_rows(A):
    """Rows whose stored off-diagonal entries are all zero (the identity rows)."""
    A = A.tocsr()
    d = A.diagonal()
    return (d != 0) & (np.abs(A).sum(axis=1).A1 == np.abs(d))


def strength(A, free, theta):
    """Strong couplings (CSR pattern, no diagonal) between free nodes and their scaled
    magnitude |a_ij| / sqrt(a_ii a_jj)."""
    A = A.tocoo()
    d = np.sqrt(np.abs(A.tocsr().diagonal()))
    keep = (A.row != A.col) & free[A.row] & free[A.col]
    r, c, v = A.row[keep], A.col[keep], A.data[keep]
    s = np.abs(v) / (d[r] * d[c])
    strong = s > max(theta, 1e-10)
    S = sp.csr_matrix((s[strong], (r[strong], c[strong])), shape=A.shape)
    S.sort_indices()
    return S


def aggregate(S, free):
    """Aggregate index per node (-1 for the nodes left out) and the aggregate count."""
    n = S.shape[0]
    ptr, idx, val = S.indptr, S.indices, S.data
    agg = np.full(n, -1, dtype=np.int64)
    na = 0
    for i in range(n):                                   # pass 1
        if not free[i] or agg[i] >= 0:
            continue
        nb = idx[ptr[i]:ptr[i + 1]]
        if np.all(agg[nb] < 0):
            agg[i] = na
            agg[nb] = na
            na += 1
    first = agg.copy()
    for i in range(n):                                   # pass 2
        if not free[i] or agg[i] >= 0:
            continue
        nb, sv = idx[ptr[i]:ptr[i + 1]], val[ptr[i]:ptr[i + 1]]
        best, pick = -1.0, -1
        for j, s in zip(nb, sv):                         # ascending j: ties to the lower index
            if first[j] >= 0 and s > best * (1 + 1e-9):
                best, pick = s, j
        if pick >= 0:
            agg[i] = first[pick]
    for i in range(n):                                   # pass 3
        if not free[i] or agg[i] >= 0:
            continue
        agg[i] = na
        for j in idx[ptr[i]:ptr[i + 1]]:
            if agg[j] < 0:
                agg[j] = na
        na += 1
    return agg, na


def gershgorin(A):
    """max_i sum_j |a_ij| / a_ii."""
    A = A.tocsr()
    return float(np.max(np.abs(A).sum(axis=1).A1 / A.diagonal()))


def omegas(deg, lo, hi):
    """Reciprocal Chebyshev nodes of [lo, hi], largest node first."""
    k = np.arange(deg)
    return 1.0 / ((hi + lo) / 2 + (hi - lo) / 2 * np.cos(np.pi * (2 * k + 1) / (2 * deg)))


class AMG:
    def __init__(self, A, theta=0.02, coarse=500, max_levels=10, deg=2, alpha=30.0, deg_coarse=1):
        A = sp.csr_matrix(A, dtype=np.float64)
        A.sort_indices()
        free = ~dirichlet_rows(A)
        self.levels = []
        while True:
            n = A.shape[0]
            rho = gershgorin(A)
            d = deg if not self.levels else deg_coarse
            lev = dict(A=A, dinv=1.0 / A.diagonal(), rho=rho, om=omegas(d, rho / alpha, rho))
            self.levels.append(lev)
            if n <= coarse or len(self.levels) == max_levels:
                break
            S = strength(A, free, theta)
            agg, na = aggregate(S, free)
            if na > 0.8 * n:
                break
            rows = np.flatnonzero(agg >= 0)
            T = sp.csr_matrix((np.ones(len(rows)), (rows, agg[rows])), shape=(n, na))
            w = 4.0 / (3.0 * rho)
            P = (T - w * sp.diags(lev["dinv"]) @ (A @ T)).tocsr()
            P.eliminate_zeros()
            P.sort_indices()
            lev["P"], lev["agg"], lev["T"] = P, agg, T
            A = (P.T @ (A @ P)).tocsr()
            A.eliminate_zeros()
            A.sort_indices()
            free = np.ones(A.shape[0], dtype=bool)
        last = self.levels[-1]
        last["inv"] = np.linalg.inv(last["A"].toarray())

    def sizes(self):
        return [lev["A"].shape[0] for lev in self.levels]

    def apply(self, r, l=0):
        lev = self.levels[l]
        if "P" not in lev:
            return lev["inv"] @ r
        A, di, om = lev["A"], lev["dinv"], lev["om"]
        x = om[0] * di * r
        for w in om[1:]:
            x = x + w * di * (r - A @ x)
        xc = self.apply(lev["P"].T @ (r - A @ x), l + 1)
        x = x + lev["P"] @ xc
        for w in om[::-1]:
            x = x + w * di * (r - A @ x)
        return x
