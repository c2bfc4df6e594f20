"""Oracle nonlinear CG with line search (FP64).  Test infrastructure only.

Follows ``scripts/minimizers.py`` and ``scripts/jax_tools.py``:

* ``compute_h`` 16-31: ``min(2 sqrt(eps)(1+|x|_2)/|d|_2, sqrt(eps)/|d|_inf)``.
* ``alpha_init`` 34-57 (c = 0.1, o = 0.99).
* ``line_search`` 92-137 (rho 0.5, c 1e-4, delta 0.1, eps 1e-6, 20 halvings).
* ``hestenes_stiefel_ncg`` 139-190 (HS+, restart tests with the *old* g).
* ``update_m`` / ``normalize_vectors`` jax_tools.py:4-18 (zero norms divide by 1).

``func(x) -> (f, g, alt)`` closes over the warm-start state; ``M_inv(g) -> z``.
"""
import numpy as np

EPS = np.finfo(np.float64).eps
FMAX = np.finfo(np.float64).max


def normalize_vectors(m):
    v = m.reshape(-1, 3)
    n = np.linalg.norm(v, axis=1, keepdims=True)
    return (v / np.where(n == 0, 1.0, n)).ravel()


def update_m(m, step, d):
    return normalize_vectors(m + step * d)


def compute_h(x, d):
    sq = np.sqrt(EPS)
    t1 = 2.0 * sq * (1.0 + np.linalg.norm(x)) / np.linalg.norm(d)
    t2 = sq / np.linalg.norm(d, ord=np.inf)
    return t1 if t1 < t2 else t2


def alpha_init(f0, f, g, d, g1, h, c=0.1, o=0.99):
    gd = float(np.dot(g, d))
    term1 = -h * (gd / (float(np.dot(g1, d)) - gd))
    if term1 < 0:
        term1 = c * abs(term1)
    with np.errstate(over="ignore"):
        term2 = 2.0 * (np.float64(f) - np.float64(f0)) / gd
    if term2 < 0:
        term2 = 1.0e30
    return term1 if term1 < o * term2 else term2


def line_search(x, f0, f, g, d, func, update_x=update_m, stats=None):
    rho, c, delta, eps, max_iter = 0.5, 1e-4, 0.1, 1e-6, 20
    h = compute_h(x, d)
    x1 = update_x(x, h, d)
    f1, g1 = func(x1)
    alpha = alpha_init(f0, f, g, d, g1, h)
    gd = float(np.dot(g, d))
    x1 = update_x(x, alpha, d)
    f1, g1 = func(x1)
    gd1 = float(np.dot(g1, d))
    i = 0
    while True:
        eps_k = eps * abs(f)
        wolfe = ((2 * delta - 1) * gd >= gd1) and (f1 <= f + eps_k)
        done = wolfe or (f1 <= f + c * alpha * gd) or (i > max_iter)
        if done:
            break
        alpha *= rho
        x1 = update_x(x, alpha, d)
        f1, g1 = func(x1)
        gd1 = float(np.dot(g1, d))
        i += 1
    if stats is not None:
        stats["backtracks"] = stats.get("backtracks", 0) + i
    return x1, f1, g1


def hestenes_stiefel_ncg(x, func, M_inv, tol, update_x=update_m, iter_max=100000, stats=None):
    """Returns (f, x, k).  ``stats`` (optional dict) counts the line-search halvings
    (``backtracks``) and the two descent restarts: to -g before the line search
    (``restarts_g``, minimizers.py:171-173) and to -z1 after the HS update
    (``restarts_z``, minimizers.py:182-184)."""
    st = stats if stats is not None else {}
    f0 = FMAX
    f, g = func(x)
    d = -M_inv(g)
    k = 1
    x0 = x
    while True:
        a = (f0 - f) > tol * (1 + abs(f))
        b = (np.linalg.norm(x0 - x, ord=np.inf) > np.sqrt(tol) * (1 + np.linalg.norm(x, ord=np.inf))) or k == 0
        cc = np.linalg.norm(g, ord=np.inf) > np.cbrt(tol) * (1 + abs(f))
        if not ((a or b or cc) and k < iter_max):
            break
        if np.dot(d, g) > -0.001 * np.linalg.norm(d) * np.linalg.norm(g):
            d = -g
            st["restarts_g"] = st.get("restarts_g", 0) + 1
        x1, f1, g1 = line_search(x, f0, f, g, d, func, update_x, st)
        z1 = M_inv(g1)
        y = g1 - g
        beta = max(float(np.dot(y, z1)) / float(np.dot(y, d)), 0.0)
        d = -z1 + beta * d
        if np.dot(d, g) > -0.001 * np.linalg.norm(d) * np.linalg.norm(g):
            d = -z1
            st["restarts_z"] = st.get("restarts_z", 0) + 1
        k, x0, f0, x, f, g = k + 1, x, f, x1, f1, g1
    return f, x, k
