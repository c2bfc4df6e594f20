"""Host-side mirror of the reference's set-up and function protocol.

``Model`` replaces ``escript2arrays`` (scripts/mapping.py:171-236): it reads the mesh
and materials, applies escript's boundary rules (scripts/matrix.py:49-70,
scripts/trafo.py:4-16) and builds the device operators through the C ABI.  Its
methods are the reference's hot-path functions on device vectors (torch tensors on
``cuda``): ``total_eg`` (energies.py:67-99), ``M_inv`` (loop.py:40-55), ``update_m``
(jax_tools.py:15-18), ``minimize`` (loop.py:58-65), ``compute_mh`` (loop.py:25-38).
"""
import ctypes
import math

import numpy as np

from . import _lib
from .fileio import read_fly, read_krn

SQRT_EPS = math.sqrt(np.finfo(np.float64).eps)


def _where_zero(arg):
    """escript whereZero with tol=None: |arg| <= sqrt(eps) * Lsup(arg)."""
    return np.abs(arg) <= SQRT_EPS * np.max(np.abs(arg))


def boundary_setup(coords, tets, elem_tags, tags):
    """Dirichlet mask, shell flags and radii exactly as scripts/matrix.py:49-70 decides.

    Airbox (exactly 8 nodes at max |x|): u = 0 on the 6 faces x_c = +-sup(x_c)
    (trafo.py:4-16), tensor k = I.  Otherwise: u = 0 where |x| = Rinf and the last tag
    carries the Imhoff shell tensor with R = max |x| over nodes of tag[-2] elements.
    """
    r = np.linalg.norm(coords, axis=1)
    Rinf = float(r.max())
    bmask = _where_zero(r - Rinf)
    airbox = abs(int(bmask.sum()) - 8) < 1e-8
    if airbox:
        sup = coords.max(axis=0)
        dmask = np.zeros(len(coords), dtype=bool)
        for a in range(3):
            dmask |= _where_zero(coords[:, a] + sup[a]) | _where_zero(coords[:, a] - sup[a])
        return dict(airbox=True, dirichlet=dmask, shell=None, R=0.0, Rinf=Rinf)
    if len(tags) < 2:
        raise ValueError("shell transform needs at least two element tags")
    touch = np.zeros(len(coords), dtype=bool)
    touch[tets[elem_tags == tags[-2]].ravel()] = True
    R = float((r * touch).max())
    shell = elem_tags == tags[-1]
    return dict(airbox=False, dirichlet=bmask, shell=shell, R=R, Rinf=Rinf)



def initial_state(coords, magnet, v, state_id):
    """getM (scripts/magnetization.py:54-71): the state times the magnet mask
    (wherePositive(meas), mapping.py:196), then escript's normalize (zero vectors stay 0).

    state_id 0 uniform v; 1 flower (getFlower, :24-33); 2 vortex about the x axis
    (getVortex, :13-21); 3 twisted (getTwisted, :36-51).  Points where a formula divides
    0 by 0 (the vortex's or the twisted state's axis) take the limit 0 for that term.
    """
    x = np.asarray(coords, dtype=np.float64)
    n = len(x)
    if state_id == 4:
        raise ValueError("state 'random' uses escript's RandomData generator, which has no "
                         "reproducible counterpart here")
    m = np.zeros((n, 3))
    if state_id == 0:
        m[:] = v
    with np.errstate(divide="ignore", invalid="ignore"):
        if state_id == 2:  # vortex
            r = np.sqrt(x[:, 1] ** 2 + x[:, 2] ** 2)
            rc = 0.14 * r.max()
            m[:, 0] = np.exp(-2 * (r / rc))
            m[:, 1] = -(x[:, 2] / r) * np.sqrt(1 - np.exp(-4 * (r * r) / (rc * rc)))
            m[:, 2] = (x[:, 1] / r) * np.sqrt(1 - np.exp(-4 * (r * r) / (rc * rc)))
        elif state_id == 1:  # flower
            s = x.max()
            m[:, 0] = (x[:, 0] * x[:, 2]) / (10 * s)
            m[:, 1] = (x[:, 1] * x[:, 2]) / (10 * s)
            m[:, 2] = 1.0
        elif state_id == 3:  # twisted
            r = np.sqrt(x[:, 1] ** 2 + x[:, 0] ** 2)
            s = x.max()
            pz, nz = (x[:, 2] > 0).astype(float), (x[:, 2] < 0).astype(float)
            m[:, 0] = (x[:, 0] * x[:, 2]) / (10 * s) + 4 * (
                pz * (-x[:, 1] / r) * (x[:, 2] / s) + nz * (x[:, 1] / r) * (-x[:, 2] / s))
            m[:, 1] = (x[:, 1] * x[:, 2]) / (10 * s) + 4 * (
                pz * (x[:, 0] / r) * (x[:, 2] / s) + nz * (-x[:, 0] / r) * (-x[:, 2] / s))
            m[:, 2] = 1.0
    m = np.nan_to_num(m)
    m[~np.asarray(magnet, dtype=bool)] = 0.0
    ln = np.linalg.norm(m, axis=1)
    z = (ln == 0).astype(np.float64)
    return (m * ((1.0 - z) / (ln + z))[:, None]).ravel()

def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.MumagError("no GPU visible: the HIP hot path needs a ROCm device")
    return torch


def _ptr(t, n=None, dtype="float64"):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a CUDA (ROCm) torch tensor")
    if str(t.dtype) != f"torch.{dtype}" or not t.is_contiguous():
        raise TypeError(f"expected a contiguous {dtype} tensor, got {t.dtype}")
    if n is not None and t.numel() != n:
        raise ValueError(f"expected {n} elements, got {t.numel()}")
    return ctypes.c_void_p(t.data_ptr())


class Model:
    """Operators of one mesh + material set on the GPU."""

    def __init__(self, coords, tets, elem_tags, krn, precision="fp64", device=0, drop_tol=1e-12,
                 pcg_batch=0, length_scale=1.0e-9, grid=None, inner_tol=0.0, eval_precision="fp64"):
        """``grid`` (optional): dict(xs, ys, zs, mag0, magn) of a structured Kuhn-grid mesh
        written by ``generate``; the library then validates it and uses its stencil kernels.
        ``eval_precision`` ("fp64" or "mixed"): with precision "mixed", the arithmetic of the
        energy evaluation's operator applications (the inner solves are FP32 either way)."""
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.tets = np.ascontiguousarray(tets, dtype=np.int32)
        self.elem_tags = np.ascontiguousarray(elem_tags, dtype=np.int32)
        self.N, self.NE = len(self.coords), len(self.tets)
        self.tags = sorted(int(t) for t in np.unique(self.elem_tags))
        self.krn = krn
        self.elem_mat = np.searchsorted(self.tags, self.elem_tags).astype(np.int32)
        nm = len(self.tags)
        Js = np.zeros(nm)
        A = np.zeros(nm)
        K1 = np.zeros(nm)
        u = np.zeros((nm, 3))
        for i, t in enumerate(self.tags):
            m = krn[t]
            if m["Js"] > 0:  # scripts/materials.py:28
                Js[i], A[i], K1[i] = m["Js"], m["A"], m["K1"]
                th, ph = m["theta"], m["phi"]
                u[i] = [math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)]
        self.mat_Js, self.mat_A, self.mat_K1, self.mat_u = Js, A, K1, u
        self.bnd = boundary_setup(self.coords, self.tets, self.elem_tags, self.tags)
        self.precision = precision
        lib = _lib.load()
        dirichlet = np.ascontiguousarray(self.bnd["dirichlet"], dtype=np.uint8)
        shell = None if self.bnd["shell"] is None else np.ascontiguousarray(self.bnd["shell"], dtype=np.uint8)
        keep = [self.coords, self.tets, self.elem_mat, Js, A, K1, u, dirichlet, shell]
        desc = _lib.MeshDesc()
        desc.n_nodes, desc.n_elems = self.N, self.NE
        desc.coords = self.coords.ctypes.data_as(_lib.c_double_p)
        desc.tets = self.tets.ctypes.data_as(_lib.c_int32_p)
        desc.elem_mat = self.elem_mat.ctypes.data_as(_lib.c_int32_p)
        desc.n_mats = nm
        desc.mat_Js = Js.ctypes.data_as(_lib.c_double_p)
        desc.mat_A = A.ctypes.data_as(_lib.c_double_p)
        desc.mat_K1 = K1.ctypes.data_as(_lib.c_double_p)
        uu = np.ascontiguousarray(u)
        keep.append(uu)
        desc.mat_u = uu.ctypes.data_as(_lib.c_double_p)
        desc.dirichlet = dirichlet.ctypes.data_as(_lib.c_uint8_p)
        desc.shell = shell.ctypes.data_as(_lib.c_uint8_p) if shell is not None else None
        desc.shell_R, desc.shell_Rinf = self.bnd["R"], self.bnd["Rinf"]
        desc.length_scale = length_scale
        if grid is not None:
            gx, gy, gz = (np.ascontiguousarray(grid[k], dtype=np.float64) for k in ("xs", "ys", "zs"))
            keep += [gx, gy, gz]
            desc.grid_n[0], desc.grid_n[1], desc.grid_n[2] = len(gx) - 1, len(gy) - 1, len(gz) - 1
            desc.grid_x = gx.ctypes.data_as(_lib.c_double_p)
            desc.grid_y = gy.ctypes.data_as(_lib.c_double_p)
            desc.grid_z = gz.ctypes.data_as(_lib.c_double_p)
            for a in range(3):
                desc.grid_mag0[a] = int(grid["mag0"][a])
            desc.grid_magn = int(grid["magn"])
        self.grid = grid
        opts = _lib.Options()
        opts.precision = {"fp64": _lib.MM_PREC_FP64, "mixed": _lib.MM_PREC_MIXED}[precision]
        opts.device = device
        opts.drop_tol = drop_tol
        opts.pcg_batch = pcg_batch
        opts.use_structured = 1 if grid is not None else 0
        opts.inner_tol = inner_tol
        opts.eval_precision = {"fp64": _lib.MM_PREC_FP64, "mixed": _lib.MM_PREC_MIXED}[eval_precision]
        self._ctx = ctypes.c_void_p()
        _lib.check(lib.mm_ctx_create(ctypes.byref(desc), ctypes.byref(opts), ctypes.byref(self._ctx)),
                   "mm_ctx_create")
        del keep
        self._lib = lib
        info = _lib.Info()
        _lib.check(lib.mm_ctx_info(self._ctx, ctypes.byref(info)), "mm_ctx_info")
        self.info = {k: getattr(info, k) for k, _ in _lib.Info._fields_}
        self.V = info.volume
        self.meas = np.empty(self.N)
        _lib.check(lib.mm_get_meas_host(self._ctx, self.meas.ctypes.data_as(_lib.c_double_p)), "meas")
        self.device = device

    # -------------------------------------------------------------- construction
    @classmethod
    def from_structured(cls, mesh, structured=True, **kw):
        """From a ``generate.StructuredMesh``; ``structured=False`` forces the general path
        (a ``generate.SphereMesh`` always takes it: its nodes are not a rectilinear grid)."""
        grid = None
        if structured and not getattr(mesh, "spherical", False):
            grid = dict(xs=mesh.xs, ys=mesh.ys, zs=mesh.zs, mag0=[b[0] for b in mesh.magnet_box],
                        magn=mesh.magnet_box[0][1])
        return cls(mesh.coords(), mesh.tets(), mesh.elem_tags(), mesh.krn(), grid=grid, **kw)

    @classmethod
    def from_files(cls, base, **kw):
        coords, tets, etags = read_fly(base + ".fly")
        tags = sorted(int(t) for t in np.unique(etags))
        return cls(coords, tets, etags, read_krn(base + ".krn", tags), **kw)

    def _attach(self, net):
        """Networks created on this context are destroyed before it."""
        if not hasattr(self, "_nets"):
            self._nets = []
        self._nets.append(net)

    def close(self):
        for net in getattr(self, "_nets", []):
            net.close()
        self._nets = []
        if getattr(self, "_ctx", None) and self._ctx.value:
            self._lib.mm_ctx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------------------- initial states
    def magnet_mask(self):
        return self.meas > 0

    def uniform_m(self, v):
        """getM state 0 (scripts/magnetization.py:54-71): mask * v, escript normalize."""
        m = np.zeros((self.N, 3))
        m[self.magnet_mask()] = v
        ln = np.linalg.norm(m, axis=1)
        z = (ln == 0).astype(np.float64)
        return (m * ((1.0 - z) / (ln + z))[:, None]).ravel()

    def state_m(self, v, state_id):
        """Initial states of scripts/magnetization.py:54-71 getM (uniform, flower, vortex, twisted)."""
        return initial_state(self.coords, self.magnet_mask(), v, state_id)

    def vortex_xm(self):
        """xM of scripts/magnetization.py:73-82 with scale = V**(1/3)."""
        scale = self.V ** 0.3333333333333333
        k = 0.5 * math.pi / scale
        kr = k * (self.coords[:, 0] + self.coords[:, 1])
        m = np.stack([np.sin(kr), np.cos(kr), np.zeros(self.N)], axis=1)
        m[~self.magnet_mask()] = 0.0
        return m.ravel(), k

    # -------------------------------------------------------------- device helpers
    def empty3(self):
        return _torch().zeros(3 * self.N, dtype=_torch().float64, device=f"cuda:{self.device}")

    def empty1(self):
        return _torch().zeros(self.N, dtype=_torch().float64, device=f"cuda:{self.device}")

    def to_device(self, a):
        return _torch().tensor(np.ascontiguousarray(a, dtype=np.float64), device=f"cuda:{self.device}")

    def _sync(self):
        _torch().cuda.synchronize(self.device)

    # -------------------------------------------------------------- hot-path functions
    def external_eg(self, m, hext, hdir, g=None):
        self._sync()
        E = ctypes.c_double()
        _lib.check(self._lib.mm_external_eg(self._ctx, _ptr(m, 3 * self.N), hext, _lib.D3(*hdir),
                                            ctypes.byref(E), None if g is None else _ptr(g, 3 * self.N)),
                   "mm_external_eg")
        return E.value

    def exani_eg(self, m, g=None):
        g = self.empty3() if g is None else g
        self._sync()
        E = ctypes.c_double()
        _lib.check(self._lib.mm_exani_eg(self._ctx, _ptr(m, 3 * self.N), ctypes.byref(E), _ptr(g, 3 * self.N)),
                   "mm_exani_eg")
        return E.value, g

    def hmag_eg(self, m, u, tol, g=None):
        g = self.empty3() if g is None else g
        self._sync()
        E = ctypes.c_double()
        it = ctypes.c_int32()
        _lib.check(self._lib.mm_hmag_eg(self._ctx, _ptr(m, 3 * self.N), _ptr(u, self.N), tol,
                                        ctypes.byref(E), _ptr(g, 3 * self.N), ctypes.byref(it)),
                   "mm_hmag_eg")
        return E.value, g, it.value

    def total_eg(self, m, hext, hdir, tol_u, hmag_on, u, g_proj=None, g_full=None):
        """Returns (E, P(m, g), g_full, pcg_iterations); ``u`` is updated in place."""
        g_proj = self.empty3() if g_proj is None else g_proj
        g_full = self.empty3() if g_full is None else g_full
        self._sync()
        E = ctypes.c_double()
        it = ctypes.c_int32()
        _lib.check(self._lib.mm_total_eg(self._ctx, _ptr(m, 3 * self.N), hext, _lib.D3(*hdir), tol_u,
                                         int(hmag_on), _ptr(u, self.N), ctypes.byref(E),
                                         _ptr(g_proj, 3 * self.N), _ptr(g_full, 3 * self.N),
                                         ctypes.byref(it)), "mm_total_eg")
        return E.value, g_proj, g_full, it.value

    def pcg(self, b, x, tol, atol=0.0, maxiter=10000):
        self._sync()
        it = ctypes.c_int32()
        rc = _lib.check(self._lib.mm_pcg(self._ctx, _ptr(b, self.N), _ptr(x, self.N), tol, atol,
                                         maxiter, ctypes.byref(it)), "mm_pcg")
        return it.value, rc

    def M_inv(self, m_ref, g_full, g, max_iter, z=None):
        z = self.empty3() if z is None else z
        self._sync()
        it = ctypes.c_int32()
        _lib.check(self._lib.mm_precond(self._ctx, _ptr(m_ref, 3 * self.N), _ptr(g_full, 3 * self.N),
                                        _ptr(g, 3 * self.N), int(max_iter), _ptr(z, 3 * self.N),
                                        ctypes.byref(it)), "mm_precond")
        return z, it.value

    def update_m(self, x, alpha, d, out=None):
        out = self.empty3() if out is None else out
        self._sync()
        _lib.check(self._lib.mm_update_m(self._ctx, _ptr(x, 3 * self.N), float(alpha),
                                         _ptr(d, 3 * self.N), _ptr(out, 3 * self.N)), "mm_update_m")
        return out

    def dot(self, a, b):
        self._sync()
        out = ctypes.c_double()
        _lib.check(self._lib.mm_dot(self._ctx, _ptr(a), _ptr(b), a.numel(), ctypes.byref(out)), "mm_dot")
        return out.value

    def compute_mh(self, m, hdir):
        self._sync()
        out = ctypes.c_double()
        _lib.check(self._lib.mm_compute_mh(self._ctx, _ptr(m, 3 * self.N), _lib.D3(*hdir),
                                           ctypes.byref(out)), "mm_compute_mh")
        return out.value

    def minimize(self, m, hext, hdir, params, m_ref, u, gF, stats=None):
        """One field step (loop.py:58-65); m, u, gF updated in place; returns E."""
        P = _lib.MinParams()
        P.tol_fun, P.tol_u = params["tol_fun"], params["tol_u"]
        P.precond_iter, P.hmag_on = int(params["precond_iter"]), int(params["hmag_on"])
        P.iter_max = int(params.get("ncg_iter_max", 100000))
        st = stats if stats is not None else _lib.Stats()
        self._sync()
        E = ctypes.c_double()
        _lib.check(self._lib.mm_minimize(self._ctx, _ptr(m, 3 * self.N), float(hext), _lib.D3(*hdir),
                                         ctypes.byref(P), _ptr(m_ref, 3 * self.N), _ptr(u, self.N),
                                         _ptr(gF, 3 * self.N), ctypes.byref(E), ctypes.byref(st)),
                   "mm_minimize")
        return E.value

    def apply_C(self, v, y=None):
        y = self.empty3() if y is None else y
        self._sync()
        _lib.check(self._lib.mm_apply_C(self._ctx, _ptr(v, 3 * self.N), _ptr(y, 3 * self.N)), "mm_apply_C")
        return y

    def apply_K(self, x, y=None):
        y = self.empty1() if y is None else y
        self._sync()
        _lib.check(self._lib.mm_apply_K(self._ctx, _ptr(x, self.N), _ptr(y, self.N)), "mm_apply_K")
        return y

    def divergence(self, m, b=None):
        b = self.empty1() if b is None else b
        self._sync()
        _lib.check(self._lib.mm_divergence(self._ctx, _ptr(m, 3 * self.N), _ptr(b, self.N)), "mm_divergence")
        return b

    def gradient(self, u, g=None):
        g = self.empty3() if g is None else g
        self._sync()
        _lib.check(self._lib.mm_gradient(self._ctx, _ptr(u, self.N), _ptr(g, 3 * self.N)), "mm_gradient")
        return g

    def multigrid(self, levels=8):
        """Geometric multigrid preconditioner for the FP32 inner solve (structured mixed
        contexts) with at most ``levels`` levels; 0 restores Jacobi."""
        _lib.check(self._lib.mm_multigrid(self._ctx, int(levels)), "mm_multigrid")

    def amg(self, levels=10):
        """Smoothed-aggregation AMG preconditioner for the FP32 inner solve (general mixed
        contexts) with at most ``levels`` levels; 0 restores Jacobi."""
        _lib.check(self._lib.mm_amg(self._ctx, int(levels)), "mm_amg")

    def amg_levels(self):
        """Unknowns of every AMG level, finest first."""
        nl = ctypes.c_int32()
        sizes = (ctypes.c_int64 * 10)()
        _lib.check(self._lib.mm_amg_levels(self._ctx, ctypes.byref(nl), sizes), "mm_amg_levels")
        return [int(sizes[l]) for l in range(nl.value)]

    def amg_apply(self, r, z=None):
        """z = M r with the AMG V-cycle (float32 device vectors)."""
        torch = _torch()
        z = torch.empty(self.N, dtype=torch.float32, device=r.device) if z is None else z
        self._sync()
        _lib.check(self._lib.mm_amg_apply(self._ctx, _ptr(r, self.N, "float32"), _ptr(z, self.N, "float32")),
                   "mm_amg_apply")
        return z

    def mg_levels(self):
        """Node counts per axis (nx, ny, nz) of every V-cycle level, finest first."""
        nl = ctypes.c_int32()
        dims = (ctypes.c_int32 * (3 * 8))()
        _lib.check(self._lib.mm_mg_levels(self._ctx, ctypes.byref(nl), dims), "mm_mg_levels")
        return [tuple(dims[3 * l:3 * l + 3]) for l in range(nl.value)]

    def mg_apply(self, r, z=None):
        """z = M r with the multigrid V-cycle (float32 device vectors)."""
        torch = _torch()
        z = torch.empty(self.N, dtype=torch.float32, device=r.device) if z is None else z
        self._sync()
        _lib.check(self._lib.mm_mg_apply(self._ctx, _ptr(r, self.N, "float32"), _ptr(z, self.N, "float32")),
                   "mm_mg_apply")
        return z

    def secant_start(self, on=True):
        """Start every line-search potential solve after the first from the Galerkin step
        along the change of the solution since that first one (``mm_secant_start``); the
        answer is the plain start's to within the solve tolerance."""
        _lib.check(self._lib.mm_secant_start(self._ctx, 1 if on else 0), "mm_secant_start")

    def warm_start(self, net=None):
        """Enable the learned warm start of every potential solve with ``net``
        (``learned.Network``), or disable it (None)."""
        _lib.check(self._lib.mm_warm_start(self._ctx, None if net is None else net._net,
                                           1.0 if net is None else float(net.scale)), "mm_warm_start")

    def warm_start_stats(self):
        calls = ctypes.c_int64()
        gamma = ctypes.c_double()
        _lib.check(self._lib.mm_warm_start_stats(self._ctx, ctypes.byref(calls), ctypes.byref(gamma)),
                   "mm_warm_start_stats")
        return dict(calls=calls.value, last_gamma=gamma.value)

    def warm_start_policy(self, policy="every"):
        """``"every"``: the network warm-starts every potential solve; ``"measured"``: each
        solve takes the faster form (warm-started or plain) by measured solve time."""
        code = {"every": _lib.WS_EVERY, "measured": _lib.WS_MEASURED}[policy]
        _lib.check(self._lib.mm_warm_start_policy(self._ctx, code), "mm_warm_start_policy")

    def warm_start_report(self):
        """Solves and warm-started solves since the policy was set, with the mean wall
        seconds of each form."""
        n, nw = ctypes.c_int64(), ctypes.c_int64()
        tp, tw = ctypes.c_double(), ctypes.c_double()
        _lib.check(self._lib.mm_warm_start_report(self._ctx, ctypes.byref(n), ctypes.byref(nw), ctypes.byref(tp),
                                                  ctypes.byref(tw)), "mm_warm_start_report")
        return dict(solves=n.value, warm_solves=nw.value, mean_s_plain=tp.value, mean_s_warm=tw.value)

    def lp_setup(self):
        """Workspace of the learned preconditioner (structured contexts)."""
        _lib.check(self._lib.mm_lp_setup(self._ctx), "mm_lp_setup")

    def lp_pcg(self, b, x, tol, maxiter=10000, net=None, every=1):
        """Flexible CG with the learned preconditioner ``net`` (``learned.Network``) every
        ``every``-th iteration, Jacobi otherwise (``every=0`` or ``net=None``: the Jacobi
        PCG in the same driver); returns (iterations, network applications, rc)."""
        self._sync()
        it, apps = ctypes.c_int32(), ctypes.c_int32()
        rc = _lib.check(self._lib.mm_lp_pcg(self._ctx, None if net is None else net._net,
                                            1.0 if net is None else float(net.scale), _ptr(b, self.N),
                                            _ptr(x, self.N), float(tol), int(maxiter),
                                            int(every) if net is not None else 0, ctypes.byref(it),
                                            ctypes.byref(apps)), "mm_lp_pcg")
        return it.value, apps.value, rc

    def timing(self, on=True):
        _lib.check(self._lib.mm_timing_enable(self._ctx, int(on)), "mm_timing_enable")

    def timing_get(self):
        t = _lib.KernelTiming()
        _lib.check(self._lib.mm_timing_get(self._ctx, ctypes.byref(t)), "mm_timing_get")
        return {k: getattr(t, k) for k, _ in _lib.KernelTiming._fields_}
