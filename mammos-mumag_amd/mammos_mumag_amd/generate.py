"""Synthetic structured meshes for the benchmark workloads.

A magnet cube of edge ``E`` nm is split into ``n`` voxels per edge (``h = E/n``) and
embedded in a cubic air box of edge ``3E`` centred at the origin.  The air is ``L``
graded layers per side whose widths ``h r, h r^2, ..., h r^L`` sum to ``E``.  Every
voxel of the rectilinear grid is split into 6 Kuhn tetrahedra that share the
(0,0,0)-(1,1,1) diagonal, each ordered to positive orientation.  Nodes are numbered
x-fastest: ``i + nx (j + ny k)``.

Microstructure: ``G`` seeds uniform in the cube from ``np.random.default_rng(seed)``;
each magnet voxel takes the grain of the seed nearest to its centre.  With ``G > 1`` a
grain-boundary phase replaces every voxel that has a face neighbour in another grain
(a boundary two voxels thick).  Tags: grains ``1..G``, boundary ``G+1`` (if any), air
last.  Easy axes: ``cos(theta)`` uniform in ``[cos(cone), 1]``, ``phi`` uniform in
``[0, 2 pi)``.

Because the box is centred and has exactly 8 nodes at the largest radius, the
reference's set-up takes its air-box branch (scripts/matrix.py:49-60): ``u = 0`` on
the six faces and ``k = I``.
"""
import math

import numpy as np

GRAIN = dict(Js=1.61, A=7.7e-12, K1=4.3e6)
BOUNDARY = dict(Js=0.8, A=0.77e-12, K1=0.0)
FIELD_DIR = (0.01745, 0.0, 0.99984)

# workloads: magnet edge (nm), voxels per edge, air layers per side, grains, axis cone (deg)
CONFIGS = {
    "cube40_single": dict(edge=40.0, n=16, layers=6, grains=1, cone=0.0, hstart=8.0, hfinal=-8.0, hstep=-0.1),
    "voronoi8_100nm": dict(edge=100.0, n=60, layers=11, grains=8, cone=15.0, hstart=8.0, hfinal=-8.0, hstep=-0.05),
    "voronoi64_250nm": dict(edge=250.0, n=125, layers=14, grains=64, cone=15.0, hstart=8.0, hfinal=-8.0, hstep=-0.05),
    "voronoi512_500nm": dict(edge=500.0, n=250, layers=16, grains=512, cone=90.0, hstart=6.0, hfinal=-6.0, hstep=-0.02),
}


def layer_ratio(n, L):
    """r with sum_{k=1..L} r^k = n (the air layers span exactly one magnet edge)."""
    if L <= 0:
        return 1.0
    lo, hi = 1.0, float(n)
    for _ in range(200):
        r = 0.5 * (lo + hi)
        s = sum(r ** k for k in range(1, L + 1))
        lo, hi = (r, hi) if s < n else (lo, r)
    return 0.5 * (lo + hi)


def axis_coords(edge, n, L):
    """Node coordinates along one axis: graded air, uniform magnet, graded air."""
    h = edge / n
    r = layer_ratio(n, L)
    widths = [h * r ** k for k in range(1, L + 1)]
    half = edge / 2
    left = [-half]
    for w in widths:
        left.append(left[-1] - w)
    left = left[::-1]  # from -1.5 E up to -E/2
    mag = [-half + h * i for i in range(1, n)]
    right = [half]
    for w in widths:
        right.append(right[-1] + w)
    xs = np.array(left + mag + right, dtype=np.float64)
    # make the outer faces exactly +-1.5 E and the magnet faces exactly +-E/2
    xs[0], xs[-1] = -1.5 * edge, 1.5 * edge
    xs[L], xs[L + n] = -half, half
    return xs


# the 6 Kuhn paths through the unit voxel: corner offsets (dx, dy, dz) of the 4 nodes
_PERMS = [(0, 1, 2), (1, 2, 0), (2, 0, 1), (0, 2, 1), (2, 1, 0), (1, 0, 2)]


def _kuhn_offsets():
    tets = []
    for p in _PERMS:
        c = [0, 0, 0]
        nodes = [tuple(c)]
        for ax in p:
            c[ax] = 1
            nodes.append(tuple(c))
        tets.append(nodes)
    # positive orientation: odd permutations swap nodes 1 and 2
    out = []
    for t in tets:
        a = np.array(t, dtype=float)
        det = np.linalg.det(a[1:] - a[0])
        if det < 0:
            t = [t[0], t[2], t[1], t[3]]
        out.append(t)
    return out


KUHN = _kuhn_offsets()


class StructuredMesh:
    """Arrays of a generated mesh plus its grid description."""

    def __init__(self, xs, ys, zs, voxel_tag, tags_props, magnet_box, meta):
        self.xs, self.ys, self.zs = xs, ys, zs
        self.voxel_tag = voxel_tag          # int32 [nz-1, ny-1, nx-1]
        self.tags_props = tags_props        # tag -> dict(theta, phi, K1, Js, A)
        self.magnet_box = magnet_box        # (first voxel index, voxels) per axis
        self.meta = meta

    @property
    def shape(self):
        return len(self.xs), len(self.ys), len(self.zs)

    def coords(self):
        nx, ny, nz = self.shape
        Z, Y, X = np.meshgrid(self.zs, self.ys, self.xs, indexing="ij")
        return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    def tets(self):
        """int32 [6 * voxels, 4]; voxel-major, Kuhn tet minor."""
        nx, ny, nz = self.shape
        k, j, i = np.meshgrid(np.arange(nz - 1), np.arange(ny - 1), np.arange(nx - 1), indexing="ij")
        base = (i + nx * (j + ny * k)).ravel().astype(np.int64)
        out = np.empty((base.size, 6, 4), dtype=np.int32)
        for t, nodes in enumerate(KUHN):
            for v, (dx, dy, dz) in enumerate(nodes):
                out[:, t, v] = base + dx + nx * (dy + ny * dz)
        return out.reshape(-1, 4)

    def elem_tags(self):
        return np.repeat(self.voxel_tag.ravel(), 6).astype(np.int32)

    def krn(self):
        return {t: dict(p) for t, p in self.tags_props.items()}

    def krn_rows(self):
        return [(p["theta"], p["phi"], p["K1"], 0.0, p["Js"], p["A"]) for _, p in sorted(self.tags_props.items())]


def _microstructure(edge, n, grains, cone, rng):
    """Magnet voxel tags [n, n, n] (grains 1..G, boundary G+1) and the grain properties."""
    h = edge / n
    c = -edge / 2 + h * (np.arange(n) + 0.5)
    if grains > 1:
        from scipy.spatial import cKDTree
        seeds = rng.uniform(-edge / 2, edge / 2, size=(grains, 3))
        Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
        _, lab = cKDTree(seeds).query(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1))
        lab = lab.reshape(n, n, n).astype(np.int32)
        gb = np.zeros_like(lab, dtype=bool)
        for ax in range(3):
            d = np.diff(lab, axis=ax) != 0
            sl_lo = [slice(None)] * 3
            sl_hi = [slice(None)] * 3
            sl_lo[ax] = slice(0, n - 1)
            sl_hi[ax] = slice(1, n)
            gb[tuple(sl_lo)] |= d
            gb[tuple(sl_hi)] |= d
        mag_tag = lab + 1
        mag_tag[gb] = grains + 1
    else:
        mag_tag = np.ones((n, n, n), dtype=np.int32)
    props = {}
    cmin = math.cos(math.radians(cone))
    for g in range(1, grains + 1):
        ct = rng.uniform(cmin, 1.0) if cone > 0 else 1.0
        ph = rng.uniform(0.0, 2 * math.pi) if cone > 0 else 0.0
        props[g] = dict(theta=math.acos(ct), phi=ph, **GRAIN)
    if grains > 1:
        props[grains + 1] = dict(theta=0.0, phi=0.0, **BOUNDARY)
    return mag_tag, props


AIR = dict(theta=0.0, phi=0.0, Js=0.0, A=0.0, K1=0.0)


def build(edge, n, layers, grains=1, cone=0.0, seed=1000, **_):
    xs = axis_coords(edge, n, layers)
    nxv = len(xs) - 1
    rng = np.random.default_rng(seed)
    mag_tag, props = _microstructure(edge, n, grains, cone, rng)
    air_tag = grains + 2 if grains > 1 else 2
    vt = np.full((nxv, nxv, nxv), air_tag, dtype=np.int32)
    vt[layers:layers + n, layers:layers + n, layers:layers + n] = mag_tag
    props[air_tag] = dict(AIR)
    present = set(np.unique(vt).tolist())
    props = {t: p for t, p in props.items() if t in present}
    meta = dict(edge=edge, n=n, layers=layers, grains=grains, cone=cone, seed=seed, h=edge / n,
                ratio=layer_ratio(n, layers))
    return StructuredMesh(xs, xs.copy(), xs.copy(), vt, props, ((layers, n),) * 3, meta)


def _graded(first, total, count):
    """Widths first q^k (k = 0..count-1) summing to total (q by bisection)."""
    def span(q):
        return sum(first * q ** k for k in range(count))
    lo, hi = 1e-6, 1.0
    while span(hi) < total:
        hi *= 2.0
    for _ in range(200):
        q = 0.5 * (lo + hi)
        lo, hi = (q, hi) if span(q) < total else (lo, q)
    q = 0.5 * (lo + hi)
    w = np.array([first * q ** k for k in range(count)])
    return w * (total / w.sum())


class SphereMesh(StructuredMesh):
    """A magnet cube in a sphere of free space inside a spherical transformation shell.

    The nodes are those of a box-in-box voxel grid (``xs``, ``ys``, ``zs`` are its parameter
    coordinates); ``coords`` moves every node outside the magnet along its own ray from the
    origin: a node of layer ``l`` (Chebyshev radius ``s_l`` in parameter space) to radius
    ``(1 - t) a |d| + t R`` with ``t = (s_l - a) / (R - a)`` (``a |d|``: where the ray leaves
    the magnet cube) up to layer ``T`` on the sphere of radius ``R``, then to ``s_l`` up to
    ``Rinf``.  Along every ray the radius grows with the layer, so the map is one-to-one.

    Each voxel is split into the 6 Kuhn tetrahedra of its own octant: the shared diagonal
    runs from the corner nearest the origin to the farthest one (mirror images of the
    airbox grid's split in the other octants; neighbouring voxels still share face
    diagonals).  Every tetrahedron then lies where one coordinate has the largest modulus,
    where the map is smooth, and keeps a positive volume.  Not a rectilinear grid: only the
    general (unstructured) path takes it."""

    spherical = True

    def tets(self):
        """int32 [6 * voxels, 4]; voxel-major, Kuhn tet minor, mirrored per octant."""
        nx, ny, nz = self.shape
        k, j, i = (v.ravel() for v in np.meshgrid(np.arange(nz - 1), np.arange(ny - 1), np.arange(nx - 1),
                                                  indexing="ij"))
        flip = [(0.5 * (c[1:] + c[:-1]) < 0).astype(np.int64) for c in (self.xs, self.ys, self.zs)]
        fi, fj, fk = flip[0][i], flip[1][j], flip[2][k]
        base = (i + nx * (j + ny * k)).astype(np.int64)
        out = np.empty((base.size, 6, 4), dtype=np.int64)
        for t, nodes in enumerate(KUHN):
            for v, (dx, dy, dz) in enumerate(nodes):
                out[:, t, v] = base + np.where(fi, 1 - dx, dx) + nx * (np.where(fj, 1 - dy, dy) +
                                                                       ny * np.where(fk, 1 - dz, dz))
        odd = (fi + fj + fk) % 2 == 1          # a mirror image: swap two nodes to stay positive
        out[odd, :, 1], out[odd, :, 2] = out[odd, :, 2], out[odd, :, 1].copy()
        return out.reshape(-1, 4).astype(np.int32)

    def __init__(self, xs, voxel_tag, tags_props, magnet_box, meta, s_layers, T, R, Rinf):
        super().__init__(xs, xs.copy(), xs.copy(), voxel_tag, tags_props, magnet_box, meta)
        self.s_layers, self.T, self.R, self.Rinf = s_layers, T, R, Rinf

    def coords(self):
        p = super().coords()
        a = 0.5 * self.meta["edge"]
        L = len(self.s_layers) - 1                      # air + shell layers per side
        nx, n = len(self.xs), self.meta["n"]
        idx = np.arange(nx)
        lay = np.maximum(0, np.maximum(L - idx, idx - (L + n)))   # layer of each axis index
        Z, Y, X = np.meshgrid(lay, lay, lay, indexing="ij")
        l = np.maximum(np.maximum(X, Y), Z).ravel()
        out = p.copy()
        sel = l > 0
        q, ls = p[sel], l[sel]
        s = self.s_layers[ls]                            # Chebyshev radius of the node's layer
        rn = np.linalg.norm(q, axis=1)
        dlen = rn / s                                    # |d|, d = p / |p|_inf on the unit cube
        t = np.minimum((s - a) / (self.R - a), 1.0)
        r = np.where(ls <= self.T, (1.0 - t) * a * dlen + t * self.R, s)
        r[ls == L] = self.Rinf
        r[ls == self.T] = self.R
        out[sel] = q * (r / rn)[:, None]
        return out


def build_sphere(edge, n, air_layers, shell_layers, grains=1, cone=0.0, seed=1000, R=None, Rinf=None):
    """Magnet cube of edge ``edge`` (``n`` voxels per edge, the microstructure of ``build``
    with the same seed) inside free space out to the sphere of radius ``R`` (default the
    edge, so that the cube's corners at 0.87 E stay clear of it) and a transformation shell
    from ``R`` to ``Rinf`` (default 1.8 R).  ``air_layers`` / ``shell_layers`` node layers
    per side: radial widths graded geometrically from the voxel size h along the face
    normals.  Tags: grains, boundary phase (if any), free space, shell last — the layout
    that takes the reference's shell branch (scripts/matrix.py:49-70, trafo.py:24-45):
    k = I on every tag but the last, the Imhoff tensor on the shell, u = 0 at |x| = Rinf."""
    a = 0.5 * edge
    R = float(edge if R is None else R)
    Rinf = float(1.8 * R if Rinf is None else Rinf)
    if not (R > math.sqrt(3.0) * a and Rinf > R):
        raise ValueError("need sqrt(3) edge / 2 < R < Rinf")
    h = edge / n
    w_air = _graded(h, R - a, air_layers)
    w_shell = _graded(w_air[-1], Rinf - R, shell_layers)
    s_layers = np.concatenate([[a], a + np.cumsum(w_air), R + np.cumsum(w_shell)])
    s_layers[air_layers] = R
    s_layers[-1] = Rinf
    L = air_layers + shell_layers
    mag = -a + h * np.arange(1, n)
    xs = np.concatenate([-s_layers[::-1], mag, s_layers])
    rng = np.random.default_rng(seed)
    mag_tag, props = _microstructure(edge, n, grains, cone, rng)
    air_tag = grains + 2 if grains > 1 else 2
    shell_tag = air_tag + 1
    nxv = len(xs) - 1
    v = np.arange(nxv)
    vlay = np.maximum(0, np.maximum(L - v, v + 1 - (L + n)))   # outer node layer of each voxel
    Z, Y, X = np.meshgrid(vlay, vlay, vlay, indexing="ij")
    vl = np.maximum(np.maximum(X, Y), Z)
    vt = np.where(vl > air_layers, shell_tag, air_tag).astype(np.int32)
    vt[L:L + n, L:L + n, L:L + n] = mag_tag
    props[air_tag] = dict(AIR)
    props[shell_tag] = dict(AIR)
    present = set(np.unique(vt).tolist())
    props = {t: p for t, p in props.items() if t in present}
    meta = dict(edge=edge, n=n, layers=L, air_layers=air_layers, shell_layers=shell_layers, grains=grains,
                cone=cone, seed=seed, h=h, R=R, Rinf=Rinf)
    return SphereMesh(xs, vt, props, ((L, n),) * 3, meta, s_layers, air_layers, R, Rinf)


def loop_params(cfg, mfinal=-2.0):
    c = CONFIGS[cfg]
    from .fileio import _normalize
    return dict(m=[0.0, 0.0, 1.0], state_id=0, mstep=0.4, mfinal=mfinal,
                h=_normalize(list(FIELD_DIR)), hstart=c["hstart"],
                hfinal=c["hfinal"], hstep=c["hstep"], hmag_on=1, tol_fun=1e-10, tol_u=1e-10,
                precond_iter=10, iter_max=1000, verbose=0)


def build_config(cfg, seed=1000):
    c = CONFIGS[cfg]
    return build(c["edge"], c["n"], c["layers"], c["grains"], c["cone"], seed=seed)
