"""Oracle energies and gradients (FP64).  Test infrastructure only.

Follows ``scripts/energies.py``:

* Zeeman ``external_g/eg`` 7-21: ``g[3i+a] = -(hext h_a / V) meas_i``, ``E = m.g``.
* exchange+anisotropy ``exani_eg`` 30-33: ``g = C m``, ``E = m.g / 2``.
* magnetostatics ``hmag_g/eg`` 36-59: ``b = Dx mx + Dy my + Dz my``,
  ``u = jax-CG(K, diagK, b, u0, tol)``, ``g = interleave(Gx u, Gy u, Gz u)``, ``E = m.g / 2``.
* ``projection`` 61-65: ``m x (g x m)`` per node.
* ``total_eg`` 67-99: returns ``(E, P(m, g), (u, g), calls + 1)``.
"""
import numpy as np

from .solvers import jax_cg


def external_g(model, hext, hdir):
    g = np.empty((model.N, 3))
    for a in range(3):
        g[:, a] = (-hext * hdir[a] / model.V) * model.meas
    return g.ravel()


def external_eg(model, m, hext, hdir):
    g = external_g(model, hext, hdir)
    return float(np.dot(m, g)), g


def exani_eg(model, m):
    g = model.C @ m
    return 0.5 * float(np.dot(m, g)), g


def hmag_g(model, m, u0, tol, stats=None):
    b = model.Dxyz[0] @ m[0::3] + model.Dxyz[1] @ m[1::3] + model.Dxyz[2] @ m[2::3]
    u, it = jax_cg(model.K, model.diagK, b, u0, tol)
    if stats is not None:
        stats["pcg_iters"] = stats.get("pcg_iters", 0) + it
        stats["pcg_calls"] = stats.get("pcg_calls", 0) + 1
    g = np.stack([model.Gxyz[0] @ u, model.Gxyz[1] @ u, model.Gxyz[2] @ u], axis=1).ravel()
    return g, u


def hmag_eg(model, m, u0, tol, stats=None):
    g, u = hmag_g(model, m, u0, tol, stats)
    return 0.5 * float(np.dot(m, g)), g, u


def projection(m, g):
    m_ = m.reshape(-1, 3)
    g_ = g.reshape(-1, 3)
    return np.cross(m_, np.cross(g_, m_)).ravel()


def total_eg(model, m, tol, hext, hdir, hmag_on, u0, stats=None):
    """Returns (E, P(m, g), u, g_full)."""
    energy, gradient = external_eg(model, m, hext, hdir)
    e, g = exani_eg(model, m)
    energy += e
    gradient = gradient + g
    if hmag_on > 0:
        e, g, u = hmag_eg(model, m, u0, tol, stats)
        energy += e
        gradient = gradient + g
    else:
        u = u0
    if stats is not None:
        stats["calls"] = stats.get("calls", 0) + 1
    return energy, projection(m, gradient), u, gradient
