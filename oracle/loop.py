"""Oracle hysteresis loop driver (FP64).  Test infrastructure only.

Follows ``scripts/loop.py``:

* ``compute_mh`` 25-38: ``sum_i (h . m_i) meas_i / V``.
* ``M_inv`` 40-55: truncated CG on ``H v = m0 x ((C v) x m0) - (m0 . gF)_i v_i`` with
  ``m0`` the *initial* magnetisation (a module global in the reference) and ``gF``
  the full gradient of the last energy evaluation; preconditioner ``D`` (exchange
  diagonal).  Called with ``-g``.
* ``minimize`` 58-65 -> ``hestenes_stiefel_ncg`` with ``total_eg`` and ``update_m``.
* ``solve`` 71-124 / ``loop`` 126-142: field sweep, snapshot rule, ``.dat`` record.
"""
import numpy as np

from .energies import projection, total_eg
from .minimizers import FMAX, hestenes_stiefel_ncg, update_m
from .solvers import truncated_cg


def compute_mh(model, m, hdir):
    return float(np.dot(hdir[0] * m[0::3] + hdir[1] * m[1::3] + hdir[2] * m[2::3], model.meas) / model.V)


class Minimizer:
    """State threaded through the reference's ``alt_args`` = (u, gradF)."""

    def __init__(self, model, params, m_init):
        self.model = model
        self.p = params
        self.m_init = m_init.copy()
        self.u = np.zeros(model.N)
        self.gF = np.zeros(3 * model.N)
        self.stats = {}

    def func(self, hext):
        p = self.p

        def f(x):
            E, g, u, gF = total_eg(self.model, x, p["tol_u"], hext, p["h"], p["hmag_on"], self.u, self.stats)
            self.u, self.gF = u, gF
            return E, g
        return f

    def H(self, v):
        m = self.m_init
        m_ = m.reshape(-1, 3)
        r1 = projection(m, self.model.C @ v)
        r2 = np.sum(m_ * self.gF.reshape(-1, 3), axis=1, keepdims=True) * v.reshape(-1, 3)
        return r1 - r2.ravel()

    def M_inv(self, g):
        z, it = truncated_cg(self.H, -g, self.model.Dex, max_iter=self.p["precond_iter"])
        self.stats["tcg_iters"] = self.stats.get("tcg_iters", 0) + it
        return z

    def minimize(self, m, hext):
        f, x, k = hestenes_stiefel_ncg(m, self.func(hext), self.M_inv, self.p["tol_fun"], update_m,
                                       stats=self.stats)
        self.stats["ncg_iters"] = self.stats.get("ncg_iters", 0) + k
        return f, x, k


def run_loop(model, params, m0, max_steps=None, on_save=None, on_step=None):
    """scripts/loop.py:71-142. Returns rec (list of [counter, hext, mh, E]) and the minimizer."""
    hstart, hfinal, hstep = params["hstart"], params["hfinal"], params["hstep"]
    if max_steps is None:
        max_steps = int(abs(hfinal - hstart) / abs(hstep)) + 1
    mfinal, mstep, hdir = params["mfinal"], params["mstep"], params["h"]
    mini = Minimizer(model, params, m0)
    m = m0.copy()
    mh = FMAX
    last_saved = FMAX
    counter = 0
    hext = hstart
    rec = []
    idx = 0
    while mh > mfinal and idx < max_steps:
        energy, m, _ = mini.minimize(m, hext)
        mh = compute_mh(model, m, hdir)
        if (last_saved - mh) > mstep:
            counter += 1
            last_saved = mh
            if on_save is not None:
                on_save(counter, m, mini.u)
        rec.append([counter, hext, mh, energy])
        if on_step is not None:
            on_step(idx, hext, mh, energy, mini.stats)
        hext = hext + hstep
        idx += 1
    return rec, mini
