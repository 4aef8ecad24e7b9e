"""Pins the oracle (oracle/fem.py) before it is trusted as the parity reference.

Each check is chosen so that a plausible slip in the restatement (dropped term,
wrong sign or index, transposed operand, wrong node order or quadrature) fails:
reference known-answer tests, exact rational integration, finite differences of
the energy, invariants of the Hessian and brute force on tiny inputs.
"""

import json
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla

from oracle import fem

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


# --------------------------------------------------------------- KATs ------

def test_operator_kats():
    k = GOLDEN["operator_two_triangles"]
    X = np.array(k["nodes"])
    E = np.array(k["elements"], dtype=np.int32)
    t = fem.Tri3()
    np.testing.assert_allclose(fem.op_eval(t, X, E, np.array(k["eval_input"])),
                               k["eval_expected"])
    np.testing.assert_allclose(fem.op_grad(t, X, E, X @ np.array(k["grad_coeffs"])),
                               k["grad_expected"])
    np.testing.assert_allclose(fem.integrate_per_element(t, X, E, np.ones(4)), k["areas_expected"])
    np.testing.assert_allclose(fem.integrate(t, X, E, np.ones(4)), sum(k["areas_expected"]))
    q = np.full((2, 1), k["quad_value"])
    np.testing.assert_allclose(fem.integrate_per_element(t, X, E, q), k["quad_integral_expected"])


@pytest.mark.parametrize("cls", [fem.Tri3, fem.Tet4, fem.Hex8])
def test_element_linear_fields_exact(cls):
    """tests/test_element.py:45-147 on the oracle: exact gradients of linear fields."""
    el = cls()
    X = el.ref_nodes
    E = np.arange(el.nne, dtype=np.int32)[None, :]
    d = el.dim
    rng = np.random.default_rng(0)
    a = np.array([2.0, 3.0, 4.0][:d])
    np.testing.assert_allclose(fem.op_grad(el, X, E, X @ a)[0], np.tile(a, (len(el.quad_weights), 1)),
                               atol=1e-12)
    A = rng.standard_normal((d, d))
    g = fem.op_grad(el, X, E, X @ A.T)[0]
    for q in range(g.shape[0]):
        np.testing.assert_allclose(g[q], A, atol=1e-12)
    B = rng.standard_normal((2, 2, d))
    T = np.einsum("ijk,nk->nij", B, X)
    g = fem.op_grad(el, X, E, T)[0]
    for q in range(g.shape[0]):
        np.testing.assert_allclose(g[q], B, atol=1e-12)


def test_lifter_kats():
    k = GOLDEN["lifter_fixed_periodic"]
    L = fem.Lifter(k["size"], fem.Fixed(k["fixed"]), fem.Periodic(k["periodic_dofs"], k["periodic_master"]))
    u = L.lift_from_zeros(np.array(k["u_reduced"]))
    np.testing.assert_array_equal(u, k["lift_expected"])
    np.testing.assert_array_equal(L.reduce(u), k["u_reduced"])
    k = GOLDEN["lifter_no_constraints"]
    L = fem.Lifter(k["size"])
    np.testing.assert_array_equal(L.lift_from_zeros(np.arange(4.0)), k["lift_expected"])
    for name in ("reduce_adjoint_fixed", "reduce_adjoint_periodic", "reduce_adjoint_periodic_fixed"):
        k = GOLDEN[name]
        cons = []
        if "periodic_dofs" in k:
            cons.append(fem.Periodic(k["periodic_dofs"], k["periodic_master"]))
        if "fixed" in k:
            cons.append(fem.Fixed(k["fixed"]))
        L = fem.Lifter(k["size"], *cons)
        np.testing.assert_array_equal(L.reduce_adjoint(np.array(k["r_full"])), k["expected"])


def _dense_hessian(prob, u):
    K = prob.element_stiffness(u)
    n = prob.n
    nc = prob.nc
    dofs = (prob.elements[:, :, None] * nc + np.arange(nc)).reshape(len(prob.elements), -1)
    H = np.zeros((n, n))
    for e in range(len(dofs)):
        H[np.ix_(dofs[e], dofs[e])] += K[e]
    return H


def test_sparse_kat_tri3_elasticity():
    """tests/test_sparse.py:49-80: colored Hessian == dense Hessian; primal == grad."""
    X, E = fem.unit_square_tri3(8, 8)
    prob = fem.Problem(fem.Tri3(), X, E, fem.LinearElasticity(1.0, 0.0))
    S = fem.pattern_from_mesh(E, X.shape[0], 2)
    colors = fem.distance2_colors_py(S.indptr, S.indices, S.shape[0])
    u0 = np.zeros(prob.n)
    data = fem.colored_jacfwd(lambda s: prob.hvp(u0, s), S.indptr, S.indices, colors)
    K = sps.csr_matrix((data, S.indices, S.indptr), shape=S.shape).toarray()
    # dense Hessian by brute force: H e_j for every j
    Hd = np.stack([prob.hvp(u0, np.eye(prob.n)[j]) for j in range(prob.n)], axis=1)
    np.testing.assert_allclose(K, Hd, rtol=1e-7, atol=1e-12)
    np.testing.assert_allclose(K, _dense_hessian(prob, u0), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(prob.residual(u0), np.zeros(prob.n), atol=1e-15)


# ----------------------------------------------------------- closed forms --

def _exact_hex8_le_stiffness(mu, lmbda):
    """Unit-cube Hex8 K_e by exact rational integration of products of 1-D linears."""
    signs = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
             (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]

    def I(a, i, b, j):
        """integral over [0,1]^3 of dN_a/dx_i dN_b/dx_j."""
        val = Fraction(1)
        for d in range(3):
            sa, sb = signs[a][d], signs[b][d]
            if d == i and d == j:
                val *= sa * sb
            elif d == i:
                val *= Fraction(sa, 2)
            elif d == j:
                val *= Fraction(sb, 2)
            else:
                val *= Fraction(1, 3) if sa == sb else Fraction(1, 6)
        return val

    K = np.zeros((24, 24))
    for a in range(8):
        for i in range(3):
            for b in range(8):
                for k in range(3):
                    lap = sum(I(a, j, b, j) for j in range(3))
                    v = mu * ((lap if i == k else 0) + I(a, k, b, i)) + lmbda * I(a, i, b, k)
                    K[a * 3 + i, b * 3 + k] = float(v)
    return K


def test_hex8_le_stiffness_exact():
    mu, lm = 0.3846153846153846, 0.5769230769230769
    X, E = fem.cube_hex8(1)
    prob = fem.Problem(fem.Hex8(), X, E, fem.LinearElasticity(mu, lm))
    K = prob.element_stiffness(np.zeros(prob.n))[0]
    np.testing.assert_allclose(K, _exact_hex8_le_stiffness(mu, lm), rtol=1e-13, atol=1e-14)


def test_tri3_poisson_load_is_area_over_three():
    X, E = fem.unit_square_tri3(4, 4)
    prob = fem.Problem(fem.Tri3(), X, E, fem.Poisson(1.0))
    r = prob.residual(np.zeros(prob.n))
    load = np.zeros(prob.n)
    area = fem.integrate_per_element(fem.Tri3(), X, E, np.ones(X.shape[0]))
    np.add.at(load, E, np.repeat(area[:, None] / 3, 3, axis=1))
    np.testing.assert_allclose(-r, load, rtol=1e-14)


# ------------------------------------------------------- finite differences --

CASES = [
    ("tri3", fem.Poisson(1.0)), ("tri3", fem.LinearElasticity(1.0, 0.5)),
    ("tri3", fem.NeoHookean(1.0, 2.0)),
    ("tet4", fem.Poisson(0.7)), ("tet4", fem.LinearElasticity(0.4, 0.6)),
    ("tet4", fem.NeoHookean(1.0, 10.0)),
    ("hex8", fem.Poisson(1.0)), ("hex8", fem.LinearElasticity(0.38, 0.58)),
    ("hex8", fem.NeoHookean(1.0, 10.0)),
]


def small_problem(kind, density, jitter=0.15):
    if kind == "tri3":
        X, E = fem.unit_square_tri3(3, 2)
        X = X + jitter * 0.3 * np.random.default_rng(1).uniform(-1, 1, X.shape) * \
            (~(np.isclose(X, 0) | np.isclose(X, 1))).all(axis=1, keepdims=True) / 3
        el = fem.Tri3()
    elif kind == "tet4":
        X, E = fem.box_tet4(2, 2, 1, jitter=jitter, seed=3)
        el = fem.Tet4()
    else:
        X, E = fem.cube_hex8(2)
        rng = np.random.default_rng(4)
        inner = ~(np.isclose(X, 0) | np.isclose(X, 1)).any(axis=1)
        X = X + jitter * 0.5 * rng.uniform(-1, 1, X.shape) * inner[:, None]
        el = fem.Hex8()
    return fem.Problem(el, X, E, density)


@pytest.mark.parametrize("kind,density", CASES)
def test_residual_matches_energy_differences(kind, density):
    prob = small_problem(kind, density)
    rng = np.random.default_rng(7)
    u = 0.05 * rng.standard_normal(prob.n)
    r = prob.residual(u)
    h = 1e-6
    fd = np.array([(prob.energy(u + h * e) - prob.energy(u - h * e)) / (2 * h)
                   for e in np.eye(prob.n)])
    np.testing.assert_allclose(r, fd, rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("kind,density", CASES)
def test_hvp_matches_residual_differences(kind, density):
    prob = small_problem(kind, density)
    rng = np.random.default_rng(8)
    u = 0.05 * rng.standard_normal(prob.n)
    v = rng.standard_normal(prob.n)
    h = 1e-6
    fd = (prob.residual(u + h * v) - prob.residual(u - h * v)) / (2 * h)
    np.testing.assert_allclose(prob.hvp(u, v), fd, rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("kind,density", CASES)
def test_taylor_second_order(kind, density):
    prob = small_problem(kind, density)
    rng = np.random.default_rng(9)
    u = 0.05 * rng.standard_normal(prob.n)
    v = rng.standard_normal(prob.n)
    E0, g = prob.energy(u), prob.residual(u) @ v
    Hv = prob.hvp(u, v) @ v
    errs = []
    for h in (1e-2, 5e-3, 2.5e-3):
        errs.append(abs(prob.energy(u + h * v) - E0 - h * g - 0.5 * h * h * Hv))
    # third-order remainder: halving h divides the error by ~8 (exactly 0 if quadratic)
    if errs[0] > 1e-13:
        assert errs[1] < errs[0] / 5 and errs[2] < errs[1] / 5


@pytest.mark.parametrize("kind,density", CASES)
def test_stiffness_symmetric_and_consistent(kind, density):
    prob = small_problem(kind, density)
    u = 0.05 * np.random.default_rng(10).standard_normal(prob.n)
    K = prob.element_stiffness(u)
    np.testing.assert_allclose(K, np.swapaxes(K, 1, 2), rtol=1e-12, atol=1e-14)
    H = _dense_hessian(prob, u)
    v = np.random.default_rng(11).standard_normal(prob.n)
    np.testing.assert_allclose(H @ v, prob.hvp(u, v), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(np.diag(H), prob.hessian_diag(u), rtol=1e-12, atol=1e-14)


# ------------------------------------------------------------- invariants --

def test_poisson_constant_in_kernel_and_le_rigid_modes():
    prob = small_problem("tet4", fem.Poisson(1.0))
    np.testing.assert_allclose(prob.hvp(np.zeros(prob.n), np.ones(prob.n)), 0.0, atol=1e-13)
    prob = small_problem("hex8", fem.LinearElasticity(0.4, 0.6))
    X = prob.coords
    modes = [np.tile(np.eye(3)[i], X.shape[0]) for i in range(3)]
    for axis in range(3):   # infinitesimal rotations
        w = np.zeros(3)
        w[axis] = 1.0
        modes.append(np.cross(w, X).ravel())
    for m in modes:
        np.testing.assert_allclose(prob.hvp(np.zeros(prob.n), m), 0.0, atol=1e-12)


def test_nh_at_rest_equals_linear_elasticity():
    for kind in ("tri3", "tet4", "hex8"):
        a = small_problem(kind, fem.NeoHookean(1.3, 4.0))
        b = small_problem(kind, fem.LinearElasticity(1.3, 4.0))
        v = np.random.default_rng(12).standard_normal(a.n)
        np.testing.assert_allclose(a.hvp(np.zeros(a.n), v), b.hvp(np.zeros(b.n), v), rtol=1e-12,
                                   atol=1e-13)
        np.testing.assert_allclose(a.residual(np.zeros(a.n)), 0.0, atol=1e-14)


def test_linear_newton_converges_in_one_step():
    X, E = fem.cube_hex8(3)
    prob = fem.Problem(fem.Hex8(), X, E, fem.LinearElasticity(0.4, 0.6),
                       f_ext=fem.face_traction_loads(X, 3))
    fixed = np.where(np.isclose(X[:, 0], 0.0))[0]
    L = fem.Lifter(prob.n, fem.Fixed((fixed[:, None] * 3 + np.arange(3)).ravel()))
    S = L.adapt_sparsity(fem.pattern_from_mesh(E, X.shape[0], 3))
    colors = fem.distance2_colors(S.indptr, S.indices, S.shape[0])
    u, its, cg, hist = fem.newton(prob, L, np.zeros(prob.n), S, colors, tol=1e-9, lin_tol=1e-12)
    assert its == 1 and hist[-1] <= 1e-9 * hist[0]


# ------------------------------------------------------------ brute force --

@pytest.mark.parametrize("maker,ndpn", [(lambda: fem.unit_square_tri3(3, 4), 2),
                                        (lambda: fem.box_tet4(2, 1, 2), 3),
                                        (lambda: fem.cube_hex8(2), 1)])
def test_pattern_equals_set_of_element_pairs(maker, ndpn):
    X, E = maker()
    S = fem.pattern_from_mesh(E, X.shape[0], ndpn)
    pairs = set()
    for el in E:
        dofs = [int(n) * ndpn + c for n in el for c in range(ndpn)]
        pairs.update((a, b) for a in dofs for b in dofs)
    rows = np.repeat(np.arange(S.shape[0]), np.diff(S.indptr))
    assert set(zip(rows.tolist(), S.indices.tolist())) == pairs
    for i in range(S.shape[0]):
        seg = S.indices[S.indptr[i]:S.indptr[i + 1]]
        assert np.all(np.diff(seg) > 0)


@pytest.mark.parametrize("maker,ndpn", [(lambda: fem.unit_square_tri3(5, 4), 2),
                                        (lambda: fem.box_tet4(3, 2, 2), 3),
                                        (lambda: fem.cube_hex8(3), 3)])
def test_coloring_valid_and_c_matches_python(maker, ndpn):
    X, E = maker()
    S = fem.pattern_from_mesh(E, X.shape[0], ndpn)
    cp = fem.distance2_colors_py(S.indptr, S.indices, S.shape[0])
    cc = fem.distance2_colors(S.indptr, S.indices, S.shape[0])
    np.testing.assert_array_equal(cp, cc)
    for i in range(S.shape[0]):      # distance-2: the columns of every row differ in color
        cols = S.indices[S.indptr[i]:S.indptr[i + 1]]
        assert len(set(cc[cols].tolist())) == len(cols)


def test_greedy_is_first_fit():
    """Every color c > 0 of node i is blocked by each smaller color among earlier A^2 neighbours."""
    X, E = fem.box_tet4(2, 2, 2)
    S = fem.pattern_from_mesh(E, X.shape[0], 3)
    adj = fem.distance2_adjacency(S.shape[0], S.indptr, S.indices)
    col = fem.distance2_colors(S.indptr, S.indices, S.shape[0])
    for i in range(S.shape[0]):
        earlier = {int(col[j]) for j in adj[i] if j < i}
        assert col[i] not in earlier
        assert all(c in earlier for c in range(col[i]))


def test_pcg_matches_spsolve():
    X, E = fem.cube_hex8(3)
    prob = fem.Problem(fem.Hex8(), X, E, fem.LinearElasticity(0.4, 0.6))
    fixed = np.where(np.isclose(X[:, 0], 0.0))[0]
    L = fem.Lifter(prob.n, fem.Fixed((fixed[:, None] * 3 + np.arange(3)).ravel()))
    S = L.adapt_sparsity(fem.pattern_from_mesh(E, X.shape[0], 3))
    colors = fem.distance2_colors(S.indptr, S.indices, S.shape[0])
    K = fem.hessian_csr(prob, L, np.zeros(prob.n), S, colors)
    b = np.random.default_rng(5).standard_normal(K.shape[0])
    x, it, rel = fem.pcg(K, b, tol=1e-12)
    np.testing.assert_allclose(x, spla.spsolve(K.tocsc(), b), rtol=1e-9, atol=1e-12)
    assert rel <= 1e-12
    # single-reduction form: same solution, same iteration count up to rounding
    x1, it1, rel1 = fem.pcg_single_reduction(K, b, tol=1e-12)
    np.testing.assert_allclose(x1, spla.spsolve(K.tocsc(), b), rtol=1e-9, atol=1e-12)
    assert rel1 <= 1e-12 and abs(it1 - it) <= 1
    # the first iterates coincide with standard PCG (exact-arithmetic identity)
    for k in (1, 2, 5):
        xa, _, _ = fem.pcg(K, b, tol=0.0, maxiter=k)
        xb, _, _ = fem.pcg_single_reduction(K, b, tol=0.0, maxiter=k)
        np.testing.assert_allclose(xb, xa, rtol=1e-10, atol=1e-13 * np.abs(xa).max())
    # zero right-hand side and the iteration limit
    assert fem.pcg_single_reduction(K, np.zeros_like(b))[1] == 0
    assert fem.pcg_single_reduction(K, b, tol=1e-12, maxiter=3)[1] == 3


def test_tet4_generator_matches_reference_layout():
    """tests/test_sparse_tracer.py:29-70 connectivity; positive volumes, total volume 1."""
    X, E = fem.box_tet4(3, 2, 2)
    assert E.shape == (3 * 2 * 2 * 6, 4)
    np.testing.assert_array_equal(E[0], [0, 1, 5, 17])      # n0, n1, n3, n7 of cell (0,0,0)
    w = fem.integration_weights(fem.Tet4(), X, E)
    assert np.all(w > 0)
    np.testing.assert_allclose(w.sum(), 1.0, rtol=1e-14)
    Xj, _ = fem.box_tet4(4, 4, 4, jitter=0.2, seed=0)
    assert np.all(fem.integration_weights(fem.Tet4(), Xj, fem.box_tet4(4, 4, 4)[1]) > 0)


# ------------------------------------------------------------- periodic ----

def _periodic_square(n=5, nc=2):
    """tests/test_sparse_tracer.py:162-216 layout: Fixed bottom, Periodic left <- right."""
    X, E = fem.unit_square_tri3(n, n)
    bot = np.where(X[:, 1] == 0.0)[0]
    left = np.where(X[:, 0] == 0.0)[0]
    left = left[np.argsort(X[left, 1])]
    right = np.where(X[:, 0] == 1.0)[0]
    right = right[np.argsort(X[right, 1])]
    comp = np.arange(nc)
    fixed = (bot[:, None] * nc + comp).ravel()
    slaves = (left[:, None] * nc + comp).ravel()
    masters = (right[:, None] * nc + comp).ravel()
    return X, E, fixed, slaves, masters


def test_lift_linear_is_the_derivative_of_lift():
    X, E, fixed, slaves, masters = _periodic_square()
    n = X.shape[0] * 2
    L = fem.Lifter(n, fem.Fixed(fixed, 0.3), fem.Periodic(slaves, masters))
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal(L.size_reduced), rng.standard_normal(L.size_reduced)
    d = L.lift_from_zeros(u + v) - L.lift_from_zeros(u)
    np.testing.assert_allclose(d, L.lift_linear(v), atol=1e-14)
    # reduce_adjoint is the transpose of lift_linear
    w = rng.standard_normal(n)
    assert abs(L.lift_linear(v) @ w - v @ L.reduce_adjoint(w)) < 1e-12


def test_periodic_augment_equals_pair_closure():
    """M^T S M == {(a, b): (i, j) in S, a in {i, m(i)}, b in {j, m(j)}}."""
    X, E, fixed, slaves, masters = _periodic_square(4)
    S = fem.pattern_from_mesh(E, X.shape[0], 2)
    aug = fem.Periodic(slaves, masters).augment(S).tocoo()
    m = dict(zip(slaves.tolist(), masters.tolist()))
    Sc = S.tocoo()
    want = set()
    for i, j in zip(Sc.row.tolist(), Sc.col.tolist()):
        for a in {i, m.get(i, i)}:
            for b in {j, m.get(j, j)}:
                want.add((a, b))
    assert set(zip(aug.row.tolist(), aug.col.tolist())) == want


def test_periodic_newton_equals_direct_solve():
    """Linear elasticity with a periodic pair of edges: the oracle's Newton on
    the colored reduced Hessian equals spsolve of P^T K P u = -P^T r(g)."""
    X, E, fixed, slaves, masters = _periodic_square(5)
    top = np.where(X[:, 1] == 1.0)[0] * 2 + 1
    prob = fem.Problem(fem.Tri3(), X, E, fem.LinearElasticity(0.4, 0.6))
    L = fem.Lifter(X.shape[0] * 2, fem.Fixed(fixed, 0.0), fem.Fixed(top, 0.01),
                   fem.Periodic(slaves, masters))
    S = L.adapt_sparsity(fem.pattern_from_mesh(E, X.shape[0], 2))
    colors = fem.distance2_colors_py(S.indptr, S.indices, S.shape[0])
    u, its, cg, hist = fem.newton(prob, L, L.lift_from_zeros(np.zeros(L.size_reduced)), S, colors)
    assert hist[-1] <= 1e-10 * hist[0]
    P = np.stack([L.lift_linear(e) for e in np.eye(L.size_reduced)], axis=1)
    K = _dense_hessian(prob, np.zeros(prob.n))
    g = L.lift_from_zeros(np.zeros(L.size_reduced))
    ur = np.linalg.solve(P.T @ K @ P, -P.T @ prob.residual(g))
    np.testing.assert_allclose(u, L.lift_from_zeros(ur), atol=1e-12 * np.abs(u).max() + 1e-15)
    # the periodic edges carry equal displacements
    np.testing.assert_array_equal(u[slaves], u[masters])


def test_element_dof_pattern_oracle():
    """Pinned by brute force, and equal to pattern_from_mesh for a plain nodal field."""
    X, E = fem.unit_square_tri3(3, 2)
    nd = np.arange(X.shape[0] * 2).reshape(-1, 2)
    S = fem.pattern_from_element_dofs(nd[E].reshape(E.shape[0], -1), X.shape[0] * 2)
    Sm = fem.pattern_from_mesh(E, X.shape[0], 2)
    assert (S != Sm).nnz == 0
    rng = np.random.default_rng(1)
    ed = rng.integers(-1, 9, size=(5, 4))
    S = fem.pattern_from_element_dofs(ed, 11, diagonal_dofs=[10, 3])
    want = {(a, b) for row in ed for a in row for b in row if a >= 0 and b >= 0}
    want |= {(10, 10), (3, 3)}
    Sc = S.tocoo()
    assert set(zip(Sc.row.tolist(), Sc.col.tolist())) == want


@pytest.mark.parametrize("gen", ["tri3", "tet4", "hex8"])
def test_projection_oracle_reproduces_fe_fields(gen):
    """A field that lies in the FE space (nodal values evaluated at the
    quadrature points) projects back to its nodal values."""
    # the one-point rules of Tri3/Tet4 give a singular mass matrix: use rules
    # exact for quadratics (3 points on triangles, 4 on tetrahedra)
    if gen == "tri3":
        (X, E) = fem.unit_square_tri3(3, 2)
        el = fem.Tri3(np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]]),
                      np.full(3, 1 / 6))
    elif gen == "tet4":
        (X, E) = fem.box_tet4(2, 2, 1, jitter=0.15, seed=2)
        a, b = 0.5854101966249685, 0.1381966011250105
        el = fem.Tet4(np.array([[b, b, b], [a, b, b], [b, a, b], [b, b, a]]), np.full(4, 1 / 24))
    else:
        (X, E), el = fem.cube_hex8(2), fem.Hex8()
        inner = ~(np.isclose(X, 0) | np.isclose(X, 1)).any(axis=1)
        X = X + 0.07 * np.random.default_rng(3).uniform(-1, 1, X.shape) * inner[:, None]
    rng = np.random.default_rng(4)
    nodal = rng.standard_normal((X.shape[0], 2, 2))
    f = fem.op_eval(el, X, E, nodal)
    np.testing.assert_allclose(fem.project(el, X, E, f), nodal, atol=1e-11)
    # with constraints: free values reproduced, constrained set to their values
    bnd = np.where(np.isclose(X[:, 0], 0.0))[0]
    L = fem.Lifter(X.shape[0], fem.Fixed(bnd, 0.25))
    got = fem.project(el, X, E, f[..., 0, 0], lifter=L)
    assert np.all(got[bnd] == 0.25)


def test_interpolate_oracle_linear_fields_and_location():
    """Operator.interpolate restated: linear fields are reproduced at any point,
    the containing triangle is the first one holding the point, outside -> NaN."""
    X, E = fem.unit_square_tri3(4, 3)
    el = fem.Tri3()
    rng = np.random.default_rng(6)
    pts = rng.uniform(0, 1, (40, 2))
    pts = np.vstack([pts, X[[0, 7, 19]], [[0.5, 1 / 3], [0.25, 0.2]]])   # vertices, edges
    nodal = np.stack([1.0 + 2 * X[:, 0] - 3 * X[:, 1], X[:, 0] * 0.5], axis=1)
    got = fem.interpolate(el, X, E, nodal, pts)
    want = np.stack([1.0 + 2 * pts[:, 0] - 3 * pts[:, 1], pts[:, 0] * 0.5], axis=1)
    np.testing.assert_allclose(got, want, atol=1e-12)
    # cell (0, 0) of a 4 x 3 grid: nodes 0, 1 (= x step), 4 (= y step), 5; below the
    # diagonal lies the first triangle [n0, n1, n3] = element 0, above it element 1
    h = np.array([0.25, 1 / 3])
    idx = fem.find_containing_polygons(np.array([[0.6, 0.2], [0.2, 0.6]]) * h, X[E])
    np.testing.assert_array_equal(idx, [0, 1])
    out = fem.interpolate(el, X, E, nodal, np.array([[1.5, 0.5]]))
    assert np.isnan(out).all()


def test_containing_polygons_boundary_kat():
    """tests/test_mesh.py:7-24: a point on the edge two squares share belongs to
    the first of them."""
    k = GOLDEN["containing_polygons_boundary"]
    polys, pts = np.array(k["polygons"]), np.array(k["points"])
    np.testing.assert_array_equal(fem.find_containing_polygons(pts, polys), k["expected"])
    np.testing.assert_array_equal(fem.find_containing_polygons_batched(pts, polys), k["expected"])


def test_batched_polygon_location_equals_the_loop():
    """find_containing_polygons_batched returns the same (first) polygon as the
    per-pair loop, for interior, vertex, edge-midpoint and outside points of a
    jittered mesh whose vertices and edges are shared by several triangles."""
    X, E = fem.unit_square_tri3(7, 5)
    rng = np.random.default_rng(11)
    inner = ~np.isclose(X, 0.0).any(axis=1) & ~np.isclose(X, 1.0).any(axis=1)
    X = X + inner[:, None] * rng.uniform(-0.04, 0.04, X.shape)
    mids = 0.5 * (X[E[:, 0]] + X[E[:, 1]])
    pts = np.vstack([rng.uniform(-0.05, 1.05, (120, 2)), X, mids[::3], [[0.3, 1.0], [1.0, 1.0]]])
    want = fem.find_containing_polygons(pts, X[E])
    got = fem.find_containing_polygons_batched(pts, X[E], chunk=17)
    np.testing.assert_array_equal(got, want)
    assert (want == -1).any() and (want >= 0).sum() > 150


@pytest.mark.parametrize("kind", ["tet4", "hex8"])
def test_openmp_assembly_matches_numpy_oracle(kind):
    """oracle/fem_omp.c (the CPU baseline's element kernels, OpenMP over
    elements) against Problem's residual, HVP and Hessian diagonal on jittered
    meshes, both elastic densities, 1e-12 relative; the unsupported cases raise."""
    if kind == "tet4":
        X, el = fem.box_tet4(4, 3, 5, jitter=0.2, seed=1)
        elem = fem.Tet4()
    else:
        X, el = fem.cube_hex8(4)
        rng = np.random.default_rng(3)
        inner = ~(np.isclose(X, 0) | np.isclose(X, 1)).any(1)
        X = X + 0.05 * rng.uniform(-1, 1, X.shape) * inner[:, None]
        elem = fem.Hex8()
    for dens in (fem.LinearElasticity(0.4, 0.6), fem.NeoHookean(1.0, 10.0)):
        P = fem.Problem(elem, X, el, dens)
        Q = fem.ProblemOMP(elem, X, el, dens, threads=3)
        rng = np.random.default_rng(0)
        u = 0.02 * rng.standard_normal(P.n)
        v = rng.standard_normal(P.n)
        for a, b in ((P.residual(u), Q.residual(u)), (P.hvp(u, v), Q.hvp(u, v)),
                     (P.hessian_diag(u), Q.hessian_diag(u))):
            assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    with pytest.raises(NotImplementedError):
        fem.ProblemOMP(fem.Tri3(), *fem.unit_square_tri3(2, 2), fem.Poisson(1.0))


def test_ldf_coloring_is_a_valid_non_greedy_distance2_coloring():
    """The restated largest-degree-first coloring (_coloring.py:156-186,286-300)
    is a distance-2 coloring, differs from the natural-order greedy one, and the
    validity check rejects a coloring with one color changed to a neighbour's."""
    X, el = fem.unit_square_tri3(5, 4)
    S = fem.pattern_from_mesh(el, X.shape[0], 2)
    n = S.shape[0]
    ldf = fem.largest_degree_first_distance2_colors(S.indptr, S.indices, n)
    greedy = fem.distance2_colors_py(S.indptr, S.indices, n)
    assert fem.is_distance2_coloring(S.indptr, S.indices, ldf)
    assert fem.is_distance2_coloring(S.indptr, S.indices, greedy)
    assert not np.array_equal(ldf, greedy)
    # brute force: a distance-2 coloring gives different colors to any two columns of a row
    A = S.toarray() != 0
    for r in range(n):
        cols = np.nonzero(A[r])[0]
        assert len(set(ldf[cols])) == cols.size
    bad = greedy.copy()
    bad[S.indices[S.indptr[3] + 1]] = bad[S.indices[S.indptr[3]]]
    assert not fem.is_distance2_coloring(S.indptr, S.indices, bad)


@pytest.mark.parametrize("kind", ["tri3-poisson", "tet4-le", "hex8-nh"])
def test_per_element_parameters_split_into_uniform_problems(kind):
    """Per-element density parameters (element quantities) in the oracle: the
    energy, residual, HVP and element stiffness of a two-material mesh equal the
    sums of the two uniform problems on the element subsets (energies and nodal
    vectors are sums over elements)."""
    rng = np.random.default_rng(8)
    if kind == "tri3-poisson":
        X, el = fem.unit_square_tri3(5, 4)
        elem = fem.Tri3()
        mk = lambda a: fem.Poisson(a)  # noqa: E731
        pa, pb = (2.0,), (-0.5,)
    elif kind == "tet4-le":
        X, el = fem.box_tet4(3, 3, 3, jitter=0.2, seed=2)
        elem = fem.Tet4()
        mk = fem.LinearElasticity  # noqa: E731
        pa, pb = (1.0, 10.0), (0.3, 0.7)
    else:
        X, el = fem.cube_hex8(3)
        elem = fem.Hex8()
        mk = fem.NeoHookean  # noqa: E731
        pa, pb = (1.0, 10.0), (2.5, 4.0)
    E = el.shape[0]
    inA = X[el].mean(axis=1)[:, 0] < 0.5
    params = [np.where(inA, a, b) for a, b in zip(pa, pb)]
    P = fem.Problem(elem, X, el, mk(*params))
    PA = fem.Problem(elem, X, el[inA], mk(*pa))
    PB = fem.Problem(elem, X, el[~inA], mk(*pb))
    u = 0.02 * rng.standard_normal(P.n)
    v = rng.standard_normal(P.n)
    assert np.isclose(P.energy(u), PA.energy(u) + PB.energy(u), rtol=1e-13, atol=0)
    for a, b in ((P.residual(u), PA.residual(u) + PB.residual(u)),
                 (P.hvp(u, v), PA.hvp(u, v) + PB.hvp(u, v))):
        assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()
    K = P.element_stiffness(u)
    assert np.array_equal(K[inA], PA.element_stiffness(u)) and E == K.shape[0]


def test_projection_under_periodic_constraints_reproduces_constrained_fields():
    """The oracle's projection with Fixed and Periodic constraints (the reference's
    E_free^T M P system, utils.py:225-246): a nodal field that satisfies the
    constraints (slaves equal their masters, zero at the fixed dofs), sampled at
    the quadrature points, is reproduced exactly; the matrix is not symmetric;
    with Fixed constraints only it is M[free][:, free].  (The reference's system
    leaves out the constant part of the lift: with non-zero Fixed values the
    projected field is not reproduced there, and the oracle keeps that.)"""
    X, el = fem.unit_square_tri3(5, 4)
    n = X.shape[0]
    left = np.where(np.isclose(X[:, 0], 0.0))[0]
    left = left[np.argsort(X[left, 1])]
    right = np.where(np.isclose(X[:, 0], 1.0))[0]
    right = right[np.argsort(X[right, 1])]
    bottom = np.where(np.isclose(X[:, 1], 0.0))[0]
    slaves = np.setdiff1d(left, bottom)
    masters = right[np.isin(left, slaves)]
    g = np.cos(2 * np.pi * X[:, 0]) * X[:, 1] + X[:, 1] ** 2
    g[slaves] = g[masters]
    assert np.all(g[bottom] == 0.0)
    L = fem.Lifter(n, fem.Fixed(bottom, 0.0), fem.Periodic(slaves, masters))
    rule = np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]]), np.full(3, 1 / 6)
    tri = fem.Tri3(*rule)
    gq = fem.op_eval(tri, X, el, g)
    out = fem.project(tri, X, el, gq, lifter=L)
    assert np.abs(out - g).max() <= 1e-12 * np.abs(g).max()
    # the system matrix is not symmetric
    w = fem.integration_weights(tri, X, el)
    Nq = np.stack([tri.N(xi) for xi in tri.quad_points])
    M = np.zeros((n, n))
    Me = np.einsum("eq,qa,qb->eab", w, Nq, Nq)
    for e, nodes in enumerate(el):
        M[np.ix_(nodes, nodes)] += Me[e]
    P = np.stack([L.lift_linear(e) for e in np.eye(L.size_reduced)], axis=1)
    A = M[L.free] @ P
    assert np.abs(A - A.T).max() > 1e-3 * np.abs(A).max()
    # Fixed only: the restricted mass matrix
    LF = fem.Lifter(n, fem.Fixed(bottom, 0.0))
    PF = np.stack([LF.lift_linear(e) for e in np.eye(LF.size_reduced)], axis=1)
    assert np.array_equal(M[LF.free] @ PF, M[np.ix_(LF.free, LF.free)])
