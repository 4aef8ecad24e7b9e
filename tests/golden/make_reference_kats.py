"""Writes reference_kats.json: the known-answer vectors held by tatva's own tests
for this path (inputs and expected outputs, transcribed with their file:line).

The reference cannot be imported here (no JAX, Python 3.10), so these vectors
are the literal inputs/expected values asserted in the reference test files.
Run: python tests/golden/make_reference_kats.py
"""

import json
import os

KATS = {
    "operator_two_triangles": {
        "source": "tests/test_operator.py:14-32,113-144",
        "nodes": [[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]],
        "elements": [[0, 1, 2], [0, 2, 3]],
        "eval_input": [0.0, 1.0, 2.0, 3.0],
        "eval_expected": [[1.0], [5.0 / 3.0]],
        "grad_coeffs": [2.0, 3.0],
        "grad_expected": [[[2.0, 3.0]], [[2.0, 3.0]]],
        "areas_expected": [0.5, 0.5],
        "quad_value": 4.0,
        "quad_integral_expected": [2.0, 2.0],
    },
    "interpolate_two_triangles": {
        "source": "tests/test_operator.py:146-167 (on the mesh of :14-32)",
        "nodal": "sum of the node coordinates",
        "points": [[0.25, 0.25], [0.75, 0.25], [0.25, 0.75], [0.5, 0.5]],
        "expected": [0.5, 1.0, 1.0, 1.0],
        "outside_points": [[1.5, 0.5]],
    },
    "containing_polygons_boundary": {
        "source": "tests/test_mesh.py:7-24",
        "polygons": [[[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]],
                     [[1.0, 0.0], [2.0, 0.0], [2.0, 1.0], [1.0, 1.0]]],
        "points": [[0.5, 0.5], [1.0, 0.5], [1.5, 0.5]],
        "expected": [0, 0, 1],
    },
    "lifter_fixed_periodic": {
        "source": "tests/test_lifter.py:26-41",
        "size": 6,
        "fixed": [0, 5],
        "periodic_dofs": [2],
        "periodic_master": [1],
        "u_reduced": [10.0, 20.0, 30.0],
        "lift_expected": [0.0, 10.0, 10.0, 20.0, 30.0, 0.0],
    },
    "lifter_no_constraints": {
        "source": "tests/test_lifter.py:14-23",
        "size": 4,
        "lift_expected": [0.0, 1.0, 2.0, 3.0],
    },
    "reduce_adjoint_fixed": {
        "source": "tests/test_lifter_residual.py:25-36",
        "size": 4,
        "fixed": [0, 3],
        "r_full": [10.0, 20.0, 30.0, 40.0],
        "expected": [20.0, 30.0],
    },
    "reduce_adjoint_periodic": {
        "source": "tests/test_lifter_residual.py:7-22",
        "size": 4,
        "periodic_dofs": [2],
        "periodic_master": [1],
        "r_full": [10.0, 20.0, 30.0, 40.0],
        "expected": [10.0, 50.0, 40.0],
    },
    "reduce_adjoint_periodic_fixed": {
        "source": "tests/test_lifter_residual.py:59-78",
        "size": 4,
        "periodic_dofs": [2],
        "periodic_master": [1],
        "fixed": [3],
        "r_full": [10.0, 20.0, 30.0, 40.0],
        "expected": [10.0, 50.0],
    },
    "lifted_fixed_sum": {
        "source": "tests/test_lifter.py:93-102",
        "size": 4,
        "fixed": [0, 3],
        "u_reduced": [10.0, 20.0],
        "lift_expected": [0.0, 10.0, 20.0, 0.0],
        "sum_expected": 30.0,
    },
    "sparse_tri3_elasticity": {
        "source": "tests/test_sparse.py:41-80",
        "mesh": "unit_square(8, 8), Tri3",
        "n_dofs_per_node": 2,
        "mu": 1.0,
        "lmbda": 0.0,
        "assertion": "colored sparse.jacfwd / linearized_jacfwd Hessian == dense "
                     "jax.jacfwd(jax.jacrev(E)) at u = 0 (rtol 1e-7), primal == jax.grad(E)",
    },
    "element_linear_fields": {
        "source": "tests/test_element.py:45-147",
        "scalar_coeffs_2d": [2.0, 3.0],
        "scalar_coeffs_3d": [2.0, 3.0, 4.0],
        "assertion": "gradient of a linear field on the reference nodes is exact (atol 1e-12) "
                     "for scalar, vector (grad[i,j] = A[i,j]) and tensor (grad[i,j,k] = B[i,j,k]) "
                     "fields",
    },
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kats.json")
    with open(out, "w") as f:
        json.dump(KATS, f, indent=1)
    print(out)
