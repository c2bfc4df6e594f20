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


def build(edge, n, layers, grains=1, cone=0.0, seed=1000, **_):
    xs = axis_coords(edge, n, layers)
    nxv = len(xs) - 1
    rng = np.random.default_rng(seed)
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
        air_tag = grains + 2
    else:
        mag_tag = np.ones((n, n, n), dtype=np.int32)
        air_tag = 2
    vt = np.full((nxv, nxv, nxv), air_tag, dtype=np.int32)
    vt[layers:layers + n, layers:layers + n, layers:layers + n] = mag_tag
    props = {}
    cmin = math.cos(math.radians(cone))
    for g in range(1, grains + 1):
        ct = rng.uniform(cmin, 1.0) if cone > 0 else 1.0
        ph = rng.uniform(0.0, 2 * math.pi) if cone > 0 else 0.0
        props[g] = dict(theta=math.acos(ct), phi=ph, **GRAIN)
    if grains > 1:
        props[grains + 1] = dict(theta=0.0, phi=0.0, **BOUNDARY)
    props[air_tag] = dict(theta=0.0, phi=0.0, Js=0.0, A=0.0, K1=0.0)
    present = set(np.unique(vt).tolist())
    props = {t: p for t, p in props.items() if t in present}
    meta = dict(edge=edge, n=n, layers=layers, grains=grains, cone=cone, seed=seed, h=h,
                ratio=layer_ratio(n, layers))
    return StructuredMesh(xs, xs.copy(), xs.copy(), vt, props, ((layers, n),) * 3, meta)


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
