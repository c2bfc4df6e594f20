"""FP64 CPU oracle for the mammos-mumag per-field-step hot path.

TEST INFRASTRUCTURE ONLY. Nothing in the product package may import, call or
execute this package. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

It is a from-scratch numpy/scipy restatement of the reference algorithm
(mammos-mumag v0.7.1, ``src/mammos_mumag/scripts``):

=====================  ===================================================
module                 follows
=====================  ===================================================
``oracle.io``          scripts/mesh.py (.fly via finley ReadMesh),
                       scripts/materials.py:12-41 (.krn), scripts/tools.py
                       (.p2 defaults 64-120, .dat 20-24, mu0 26-27),
                       scripts/escript_tools.py:26-31 (weipa VTU)
``oracle.assemble``    scripts/matrix.py:9-74, scripts/trafo.py:4-45,
                       scripts/escript_tools.py:13-17,
                       scripts/mapping.py:171-236,
                       scripts/magnetization.py:54-82
``oracle.energies``    scripts/energies.py:7-99
``oracle.solvers``     scripts/solvers.py:7-131 and 221-231
                       (jax.scipy.sparse.linalg.cg semantics)
``oracle.minimizers``  scripts/minimizers.py:16-190, scripts/jax_tools.py
``oracle.loop``        scripts/loop.py:25-142
=====================  ===================================================

Parity pinning (what checks each function; see ``tests/test_oracle_*.py``):

* assembly + energies: the reference fixtures ``exani/cube_vortex.csv``,
  ``exani/cube_uniform.csv``, ``external/cube.csv``, ``hmag/cube.csv`` (both
  E_field and E_gradient), ``mapping/*.csv`` (float32), plus closed forms
  (uniform anisotropy = -K1, Zeeman = -Js h.m, K 1 = 0 off the Dirichlet
  set, G = D^T / V, symmetric C and K, mass row sums = vol/4).
* JAX CG: a direct sparse solve (tol -> 0) and the iteration invariants.
* truncated CG: a dense brute-force restatement on tiny SPD / indefinite
  systems (exact CG steps, negative-curvature exit).
* NCG + line search + loop driver: ``loop/cube.dat`` and
  ``loop/cube_0001.vtu`` of the reference's own test_loop.
"""
