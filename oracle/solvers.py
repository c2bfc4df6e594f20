"""Oracle linear solvers (FP64).  Test infrastructure only.

* ``jax_cg`` restates ``scripts/solvers.py:222-231`` (jax_scipy_cg), i.e. JAX's
  ``jax.scipy.sparse.linalg.cg`` (third-party, JAX's public ``_cg_solve``) with
  the Jacobi preconditioner ``M(x) = x * (1 / D)``:

      r = b - A x0; z = M r; p = z; gamma = r.z; k = 0
      bound = max(tol^2 (b.b), atol^2)
      while r.r > bound and k < maxiter:
          q = A p; alpha = gamma / (p.q); x += alpha p; r -= alpha q
          z = M r; gamma' = r.z; p = z + (gamma'/gamma) p; gamma = gamma'; k += 1

  It returns the iterate; the reference discards convergence info.
* ``truncated_cg`` restates ``scripts/solvers.py:7-131``
  (truncated_cg_diag_precond_jit) including the ``p == 0 -> -grad_fk`` exit.
"""
import numpy as np


def jax_cg(A, D, b, x0, tol=1e-5, atol=0.0, maxiter=10000):
    invD = 1.0 / D
    bs = float(np.dot(b, b))
    bound = max(tol * tol * bs, atol * atol)
    x = x0.copy()
    r = b - A @ x
    z = r * invD
    p = z.copy()
    gamma = float(np.dot(r, z))
    k = 0
    while float(np.dot(r, r)) > bound and k < maxiter:
        q = A @ p
        alpha = gamma / float(np.dot(p, q))
        x += alpha * p
        r -= alpha * q
        z = r * invD
        gamma_new = float(np.dot(r, z))
        beta = gamma_new / gamma
        p = z + beta * p
        gamma = gamma_new
        k += 1
    return x, k


def truncated_cg(Bk, grad_fk, Dk, max_iter=100):
    """Truncated PCG of Nocedal-Wright alg. 7.1 (solvers.py:7-131).

    ``Bk`` is a callable. Returns ``(p, iterations)``.
    """
    norm_grad = float(np.linalg.norm(grad_fk))
    eps_k = min(0.5, np.sqrt(norm_grad) * norm_grad)
    p = np.zeros_like(grad_fk)
    r = grad_fk.copy()
    z = r / Dk
    d = -z
    i = 0
    done = False
    while i < max_iter and not done:
        Bd = Bk(d)
        denom = float(d @ Bd)
        neg = denom <= 0.0
        rz = float(r @ z)
        alpha = rz / denom if not neg else 0.0
        p_new = p + alpha * d
        r_new = r + alpha * Bd
        z_new = r_new / Dk
        conv = float(np.linalg.norm(r_new)) <= eps_k
        done = neg or conv
        if not neg:
            p = p_new
            beta = float(r_new @ z_new) / (rz + 1e-16) if not conv else 0.0
            d = -z_new + beta * d
            r, z = r_new, z_new
        i += 1
    if np.all(p == 0):
        return -grad_fk, i
    return p, i
