"""Oracle P1 operator assembly (FP64, scipy.sparse CSR).

Test infrastructure only (see ``oracle/__init__.py``).

Restates the escript ``LinearPDE`` operators the reference converts to
matrices in ``scripts/mapping.py:171-236``:

* ``C``  = exchange + anisotropy, 3N x 3N interleaved (``scripts/matrix.py:37-47``):
  ``(2A_e/V) int grad phi_i . grad phi_j`` on each component plus
  ``(-2K_e/V) u_a u_b int phi_i phi_j`` (consistent P1 mass ``vol(1+d_ij)/20``).
* ``D``  = diagonal of the exchange-only operator, 1 where ``meas == 0``
  (``scripts/mapping.py:207-209``).
* ``Dx, Dy, Dz`` = ``int Js phi_j d_x phi_i`` (``scripts/matrix.py:23-35``, B coefficient).
* ``Gx, Gy, Gz`` = ``int (Js/V) d_x phi_j phi_i`` = ``Dx^T / V`` (``scripts/matrix.py:9-21``).
* ``K`` = ``int grad phi_i . k grad phi_j`` with Dirichlet rows/cols replaced by
  identity (``scripts/matrix.py:49-70``; tensor ``k`` from ``scripts/trafo.py:24-45``,
  Dirichlet set from ``trafo.py:4-16`` or ``|x| = Rinf``).  The spatially varying
  shell tensor is integrated with the 4-point degree-2 Keast rule.
* ``meas_i = sum_e Js_e vol_e / 4`` (``scripts/escript_tools.py:13-17``) and
  ``V = sum vol_e over Js_e > 0`` (``scripts/materials.py:36-37``).
"""
import math

import numpy as np
import scipy.sparse as sp

from .io import MU0_SCRIPT, read_fly, read_krn, read_p2

KEAST_A = 0.5854101966249685
KEAST_B = 0.1381966011250105
SQRT_EPS = math.sqrt(np.finfo(np.float64).eps)


def where_zero(arg):
    """escript util.whereZero with tol=None: |arg| <= sqrt(eps) * Lsup(arg)."""
    tol = SQRT_EPS * np.max(np.abs(arg))
    return np.abs(arg) <= tol


def element_geometry(coords, tets):
    """Per element: vol f64[NE], barycentric gradients G f64[NE,4,3]."""
    X = coords[tets]
    J = X[:, 1:, :] - X[:, 0:1, :]
    det = np.linalg.det(J)
    vol = np.abs(det) / 6.0
    Jinv = np.linalg.inv(J)
    G = np.empty((len(tets), 4, 3))
    G[:, 1:, :] = np.transpose(Jinv, (0, 2, 1))
    G[:, 0, :] = -G[:, 1:, :].sum(axis=1)
    return vol, G


def _coo_scalar(tets, local, n):
    """Assemble per-element 4x4 blocks into an n x n CSR matrix (duplicates summed)."""
    rows = np.repeat(tets, 4, axis=1).ravel()
    cols = np.tile(tets, (1, 4)).ravel()
    return sp.csr_matrix((local.reshape(-1), (rows, cols)), shape=(n, n))


def t3d_tensor(xq, R, Rinf):
    """scripts/trafo.py:24-45 (Imhoff spherical-shell transform), at points xq[...,3]."""
    x = np.linalg.norm(xq, axis=-1)
    a2 = R * (Rinf - R)
    ax = Rinf / x
    a2x2 = a2 / (x * x)
    f = ax * (2 - ax) / (ax - 1) ** 2
    n = xq / x[..., None]
    k = -a2x2[..., None, None] * f[..., None, None] * n[..., :, None] * n[..., None, :]
    for a in range(3):
        k[..., a, a] = a2x2 * (1 + f * (1 - n[..., a] ** 2))
    return k


def stiffness_local(vol, G):
    """Element stiffness vol grad phi_k . grad phi_l of the P1 Laplacian (k = I), [NE, 4, 4]."""
    return vol[:, None, None] * np.einsum("eka,ela->ekl", G, G)


def laplacian(coords, tets):
    """Assembled P1 Laplacian with no boundary rows replaced (the air-box K before
    scripts/matrix.py:61-70 imposes the Dirichlet rows and columns)."""
    vol, G = element_geometry(coords, tets)
    return _coo_scalar(tets, stiffness_local(vol, G), len(coords))


class Model:
    """Mesh + materials + operators of one model directory entry ``name``."""

    def __init__(self, coords, tets, elem_tags, krn, V=None, poisson=True):
        """``V`` (optional) replaces the magnet volume that scales C and G, and
        ``poisson=False`` skips K: a model of a sub-block of a larger mesh then holds that
        mesh's C, D and G rows for the nodes whose elements all lie in the sub-block."""
        self.coords, self.tets, self.elem_tags = coords, tets, elem_tags
        self.N = len(coords)
        self.tags = sorted(set(int(t) for t in np.unique(elem_tags)))
        self.krn = krn
        self._V = V
        self._materials()
        self._assemble(poisson)

    @classmethod
    def from_files(cls, base):
        _, coords, _, tets, etags = read_fly(base + ".fly")
        tags = sorted(set(int(t) for t in np.unique(etags)))
        return cls(coords, tets, etags, read_krn(base + ".krn", tags))

    def _materials(self):
        """scripts/materials.py:12-41: per-element Js, A_eff, K_eff, easy axis u."""
        ne = len(self.tets)
        self.Js = np.zeros(ne)
        self.A = np.zeros(ne)
        self.K = np.zeros(ne)
        self.u = np.zeros((ne, 3))
        for tag in self.tags:
            mat = self.krn[tag]
            if mat["Js"] > 0:
                sel = self.elem_tags == tag
                th, ph = mat["theta"], mat["phi"]
                self.u[sel] = [math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)]
                self.K[sel] = MU0_SCRIPT * mat["K1"]
                self.Js[sel] = mat["Js"]
                self.A[sel] = MU0_SCRIPT * mat["A"] / (1.0e-9 * 1.0e-9)  # size = 1e-9 (scripts/materials.py:13)

    def _assemble(self, poisson=True):
        N, tets = self.N, self.tets
        vol, G = element_geometry(self.coords, tets)
        self.vol, self.G = vol, G
        mag = self.Js > 0
        self.V = float(vol[mag].sum()) if self._V is None else float(self._V)
        V = self.V
        # meas (escript_tools.py:13-17)
        self.meas = np.bincount(tets.ravel(), weights=np.repeat(self.Js * vol / 4.0, 4), minlength=N)
        # only magnetic elements carry exchange, anisotropy, Dx and Gx coefficients
        tm, Gm, vm = tets[mag], G[mag], vol[mag]
        GG = np.einsum("eka,ela->ekl", Gm, Gm)
        S = _coo_scalar(tm, (2.0 * self.A[mag] / V * vm)[:, None, None] * GG, N)
        del GG
        self.S = S
        mass = (np.ones((4, 4)) + np.eye(4)) / 20.0
        cK = -2.0 * self.K[mag] / V * vm
        um = self.u[mag]
        C = sp.kron(S, sp.identity(3), format="csr")
        for a in range(3):
            for b in range(3):
                w = cK * um[:, a] * um[:, b]
                if not np.any(w):
                    continue
                Mab = _coo_scalar(tm, w[:, None, None] * mass[None], N)
                E = sp.coo_matrix(([1.0], ([a], [b])), shape=(3, 3))
                C = C + sp.kron(Mab, E, format="csr")
        self.C = C.tocsr()
        dS = S.diagonal()
        D = np.repeat(dS, 3)
        D = np.where(np.repeat(self.meas, 3) == 0, 1.0, D)
        self.Dex = D
        # divergence / gradient operators (matrix.py:9-35)
        self.Dxyz = []
        jv = self.Js[mag] * vm / 4.0
        for a in range(3):
            loc = np.repeat(jv[:, None] * Gm[:, :, a], 4, axis=1).reshape(-1, 4, 4)
            # loc[e, k, l] = Js vol/4 d_a phi_k   (row k test, column l trial)
            self.Dxyz.append(_coo_scalar(tm, loc, N))
        self.Gxyz = [(Dm.T / V).tocsr() for Dm in self.Dxyz]
        if poisson:
            self._poisson(vol, G)

    def _poisson(self, vol, G):
        """scripts/matrix.py:49-70 and trafo.py:4-45."""
        x = self.coords
        r = np.linalg.norm(x, axis=1)
        Rinf = r.max()
        bmask = where_zero(r - Rinf)
        self.airbox = abs(int(bmask.sum()) - 8) < 1e-8
        N, tets = self.N, self.tets
        Kloc = stiffness_local(vol, G)
        if self.airbox:
            sup = x.max(axis=0)
            dmask = np.zeros(N, dtype=bool)
            for a in range(3):
                dmask |= where_zero(x[:, a] + sup[a]) | where_zero(x[:, a] - sup[a])
            self.R = None
        else:
            dmask = bmask
            # R = sup(|x| * MaskFromTag(tags[-2])): nodes touching elements of tag[-2]
            sel = self.elem_tags == self.tags[-2]
            touch = np.zeros(N, dtype=bool)
            touch[tets[sel].ravel()] = True
            R = float((r * touch).max())
            self.R = R
            shell = self.elem_tags == self.tags[-1]
            X = x[tets[shell]]
            lam = np.full((4, 4), KEAST_B)
            np.fill_diagonal(lam, KEAST_A)
            xq = np.einsum("qk,ekc->eqc", lam, X)
            kq = t3d_tensor(xq, R, Rinf)  # e,q,3,3
            Gs = G[shell]
            Kloc_s = np.einsum("eka,eqab,elb->ekl", Gs, kq, Gs) * (vol[shell] / 4.0)[:, None, None]
            Kloc[shell] = Kloc_s
        self.Rinf = Rinf
        K = _coo_scalar(tets, Kloc, N)
        # nullify Dirichlet rows/cols, unit diagonal
        keep = sp.diags((~dmask).astype(np.float64))
        K = (keep @ K @ keep + sp.diags(dmask.astype(np.float64))).tocsr()
        K.sort_indices()
        self.K = K
        self.diagK = K.diagonal().copy()
        self.dirichlet = dmask

    # ---- initial states (scripts/magnetization.py) ----
    def magnet_mask(self):
        return self.meas > 0

    def uniform_m(self, v):
        """getM(..., state 0): mask * v, then escript normalize (arg*(1-z)/(l+z))."""
        m = np.zeros((self.N, 3))
        m[self.magnet_mask()] = v
        ln = np.linalg.norm(m, axis=1)
        z = (ln == 0).astype(np.float64)
        return (m * ((1.0 - z) / (ln + z))[:, None]).ravel()

    def vortex_xm(self):
        """scripts/magnetization.py:73-82 xM with scale = V**(1/3)."""
        scale = self.V ** 0.3333333333333333
        k = 0.5 * math.pi / scale
        kr = k * (self.coords[:, 0] + self.coords[:, 1])
        m = np.stack([np.sin(kr), np.cos(kr), np.zeros(self.N)], axis=1)
        m[~self.magnet_mask()] = 0.0
        return m.ravel(), k


def load_model(base):
    """Model + parameters for ``base`` (path without extension)."""
    model = Model.from_files(base)
    try:
        params = read_p2(base + ".p2")
    except Exception:
        params = None
    return model, params
