"""Oracle: CPU restatement of tatva's element-assembly / colored-Hessian / PCG path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports, links or
runs anything under ``oracle/``.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may use it, and only as the
checker (or the timed CPU baseline), never as the thing shipped or measured.

Contents
--------
fem.py       NumPy/SciPy FP64 restatement of the reference algorithm (meshes,
             elements, Operator gather/quadrature, densities with hand-derived
             residual and Hessian-vector forms, pattern_from_mesh, Lifter/Fixed/
             Periodic, distance-2 greedy coloring, colored jacfwd, Jacobi-PCG,
             Newton with load steps).
coloring.c   Plain-C restatement of the natural-order greedy distance-2
             coloring (``tatva/sparse/_coloring.py:27-48,136-153``) for sizes the
             pure-Python loop cannot reach; built into ``oracle/_build/``.

How each function is pinned (checked in ``tests/test_oracle_*.py``):
  * reference known-answer tests: ``tests/test_operator.py:113-144``,
    ``tests/test_element.py:45-147``, ``tests/test_lifter.py:26-41``,
    ``tests/test_lifter_residual.py:7-78``, ``tests/test_sparse.py:49-80``;
  * closed forms: exact Hex8 stiffness by rational integration, Tri3 f=1 load
    = area/3 per node, patch tests, rigid-body modes in ker K;
  * finite differences: residual vs central differences of the energy,
    Hessian-vector product vs differences of residuals, Taylor decay;
  * brute force: the CSR pattern from a set of element pairs, distance-2
    validity of colors, dense Hessian by differences on tiny meshes;
  * library routines: PCG vs ``scipy.sparse.linalg.spsolve``.
Parity unpinned: the exact color vector of the external ``tatva_coloring``
package (absent here); the oracle pins the in-repo natural-order greedy
``tatva/sparse/_coloring.py:136-153`` instead.
"""

from . import fem  # noqa: F401
