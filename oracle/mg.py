"""Multigrid V-cycle of the mixed-precision form's inner solve: CPU restatement of
mammos-mumag_amd/csrc/mg.hip (FP64) for the GPU checks.  Test infrastructure only.

No counterpart in the reference (its potential solve is Jacobi-preconditioned).  The
restatement follows the kernels: per-axis spacing-aware coarsening (two adjacent
intervals merge when both are at most 3 hmin, hmin the level's smallest spacing over the
three axes; down to at most 16 nodes per axis),
the 7-point operator of each level's rectilinear grid with zero on the faces, linear
interpolation and its transpose, Chebyshev-Richardson smoothing x += w_k (b - A x) / diag
with w_k the reciprocal Chebyshev nodes of [2/8, 2] (degree 4) on the finest level and of
[2/15, 2] (degree 3) on the others (reverse order after the coarse correction), the
coarsest level a degree-12 polynomial on [2/100, 2] with its steps in Leja order.

Pinned by: its own operator against the oracle's assembled K (interior rows), its
symmetry, and V-cycle properties (exact on the coarsest level's polynomial, linearity);
tests/test_oracle_mg.py.
"""
import numpy as np
import scipy.sparse as sp


def _pad(c):
    return np.concatenate([[0.0], np.diff(c), [0.0]])


def _edge(h1p, h2p, a, b):
    l1, u1, l2, u2 = h1p[a], h1p[a + 1], h2p[b], h2p[b + 1]
    return 2 * l1 * l2 + l1 * u2 + u1 * l2 + 2 * u1 * u2


def operator(xs, ys, zs):
    """7-point operator on the interior (faces: identity rows, no couplings), node
    (i, j, k) = i + nx (j + ny k)."""
    nx, ny, nz = len(xs), len(ys), len(zs)
    hx, hy, hz = _pad(xs), _pad(ys), _pad(zs)
    K_, J_, I_ = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    face = (I_ == 0) | (J_ == 0) | (K_ == 0) | (I_ == nx - 1) | (J_ == ny - 1) | (K_ == nz - 1)
    sx = _edge(hy, hz, J_, K_)
    syy = _edge(hz, hx, K_, I_)
    szz = _edge(hx, hy, I_, J_)
    ix = np.where(hx > 0, 1.0 / (6.0 * np.where(hx > 0, hx, 1.0)), 0.0)
    iy = np.where(hy > 0, 1.0 / (6.0 * np.where(hy > 0, hy, 1.0)), 0.0)
    iz = np.where(hz > 0, 1.0 / (6.0 * np.where(hz > 0, hz, 1.0)), 0.0)
    cxm, cxp = sx * ix[I_], sx * ix[I_ + 1]
    cym, cyp = syy * iy[J_], syy * iy[J_ + 1]
    czm, czp = szz * iz[K_], szz * iz[K_ + 1]
    diag = cxm + cxp + cym + cyp + czm + czp
    idx = np.arange(nx * ny * nz).reshape(nz, ny, nx)
    rows, cols, vals = [], [], []
    inner = ~face
    for c, (dk, dj, di) in ((cxm, (0, 0, -1)), (cxp, (0, 0, 1)), (cym, (0, -1, 0)), (cyp, (0, 1, 0)),
                            (czm, (-1, 0, 0)), (czp, (1, 0, 0))):
        kk, jj, ii = K_[inner] + dk, J_[inner] + dj, I_[inner] + di
        nb = idx[kk, jj, ii]
        keep = ~face[kk, jj, ii]
        rows.append(idx[inner][keep])
        cols.append(nb[keep])
        vals.append(-c[inner][keep])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(idx.size, idx.size))
    d = np.where(face, 1.0, diag).ravel()
    return (A + sp.diags(d)).tocsr(), face.ravel()


MERGE = 3.0
COARSEST = 16


def coarsen(x, hmin):
    """Kept node indices of axis x: walking from the first node, two adjacent intervals
    merge (the middle node drops) when both are at most MERGE * hmin."""
    d = np.diff(x)
    keep, i = [0], 0
    while i < len(x) - 1:
        i += 2 if (i + 2 <= len(x) - 1 and d[i] <= MERGE * hmin and d[i + 1] <= MERGE * hmin) else 1
        keep.append(i)
    return np.asarray(keep)


def interp(xf, keep):
    xc = xf[keep]
    P = sp.lil_matrix((len(xf), len(xc)))
    c0 = 0
    for i in range(len(xf)):
        while c0 + 1 < len(keep) and keep[c0 + 1] <= i:
            c0 += 1
        if keep[c0] == i:
            P[i, c0] = 1.0
        else:
            t = (xf[i] - xc[c0]) / (xc[c0 + 1] - xc[c0])
            P[i, c0] = 1.0 - t
            P[i, c0 + 1] = t
    return P.tocsr()


def omegas(deg, lo, hi=2.0):
    """Chebyshev-Richardson weights: reciprocals of the Chebyshev nodes of [lo, hi],
    from the largest node down."""
    k = np.arange(deg)
    return 1.0 / ((hi + lo) / 2 + (hi - lo) / 2 * np.cos(np.pi * (2 * k + 1) / (2 * deg)))


def leja(om):
    """The weights in greedy Leja order of their nodes 1/w: the largest node first, then
    each time the node with the largest product of distances to those already taken
    (ties, which symmetric nodes make, to the lower index).  Same polynomial; in FP32 the
    partial products stay bounded."""
    nodes = 1.0 / np.asarray(om, float)
    order = [int(np.argmax(nodes))]
    while len(order) < len(nodes):
        best, pick = -1.0, -1
        for j in range(len(nodes)):
            if j in order:
                continue
            p = np.prod(np.abs(nodes[j] - nodes[order]))
            if p > best * (1 + 1e-9):
                best, pick = p, j
        order.append(pick)
    return np.asarray(om)[order]


class VCycle:
    def __init__(self, xs, ys, zs, max_levels=8, deg=3, cdeg=12, deg0=4):
        self.levels, self.axes = [], []
        axes = [np.asarray(xs, float), np.asarray(ys, float), np.asarray(zs, float)]
        while True:
            A, face = operator(*axes)
            lev = dict(A=A, face=face, dinv=1.0 / A.diagonal())
            self.levels.append(lev)
            self.axes.append(tuple(axes))
            if max(len(a) for a in axes) <= COARSEST or len(self.levels) == max_levels:
                break
            hmin = min(np.diff(a).min() for a in axes)
            keeps = [coarsen(a, hmin) for a in axes]
            if all(len(k) == len(a) for k, a in zip(keeps, axes)):
                break
            lev["P"] = sp.kron(sp.kron(interp(axes[2], keeps[2]), interp(axes[1], keeps[1])),
                               interp(axes[0], keeps[0])).tocsr()
            axes = [a[k] for a, k in zip(axes, keeps)]
        self.om = omegas(deg, 2.0 / 15.0)
        self.om0 = omegas(deg0, 2.0 / 8.0)   # the finest level's smoothing
        self.omc = leja(omegas(cdeg, 2.0 / 100.0))

    def _smooth(self, lev, x, b, oms):
        for w in oms:
            x = x + w * lev["dinv"] * (b - lev["A"] @ x)
            x[lev["face"]] = 0.0
        return x

    def apply(self, r, l=0):
        lev = self.levels[l]
        r = np.where(lev["face"], 0.0, r)
        if l == len(self.levels) - 1:
            return self._smooth(lev, np.zeros_like(r), r, self.omc)
        om = self.om0 if l == 0 else self.om
        x = self._smooth(lev, np.zeros_like(r), r, om)
        rc = lev["P"].T @ (r - lev["A"] @ x)
        rc[self.levels[l + 1]["face"]] = 0.0
        x = x + lev["P"] @ self.apply(rc, l + 1)
        x[lev["face"]] = 0.0
        return self._smooth(lev, x, r, om[::-1])
