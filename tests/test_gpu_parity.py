"""GPU parity: every HIP kernel of the path against the oracle on the same
seeded inputs.  FP64 results must agree to 1e-12 relative (the tolerance is
written in each assertion); CSR indices and colors must be bit-exact."""

import json
import os

import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla

import tatva_amd as ta
from oracle import fem
from tatva_amd import energies, lifter, solvers, sparse

pytestmark = pytest.mark.gpu

RTOL = 1e-12
GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))

ORACLE_EL = {ta.Tri3: fem.Tri3, ta.Tetrahedron4: fem.Tet4, ta.Hexahedron8: fem.Hex8}


def to_np(t):
    return t.detach().cpu().numpy()


def rel_err(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def make_mesh(kind, n=4, jitter=0.2):
    if kind == "tri3":
        m = ta.Mesh.unit_square(n, n + 1)
        rng = np.random.default_rng(11)
        X = m.coords.copy()
        inner = ~(np.isclose(X, 0) | np.isclose(X, 1)).any(axis=1)
        X += jitter / n * rng.uniform(-1, 1, X.shape) * inner[:, None]
        return m.set_coords(X), ta.Tri3()
    if kind == "tet4":
        return ta.Mesh.box((n, n - 1, n + 1), type="tetrahedron", jitter=jitter, seed=5), \
            ta.Tetrahedron4()
    m = ta.Mesh.unit_cube(n)
    rng = np.random.default_rng(12)
    X = m.coords.copy()
    inner = ~(np.isclose(X, 0) | np.isclose(X, 1)).any(axis=1)
    X += jitter / n * rng.uniform(-1, 1, X.shape) * inner[:, None]
    return m.set_coords(X), ta.Hexahedron8()


def oracle_density(d):
    if isinstance(d, energies.Poisson):
        return fem.Poisson(d.f)
    if isinstance(d, energies.LinearElasticity):
        return fem.LinearElasticity(d.mu, d.lmbda)
    return fem.NeoHookean(d.mu, d.lmbda)


def pair(kind, density, n=4, f_ext=False):
    # "hex8-affine": un-jittered cube, every element a parallelepiped
    m, el = make_mesh("hex8", n, jitter=0.0) if kind == "hex8-affine" else make_mesh(kind, n)
    op = ta.Operator(m, el)
    prob = fem.Problem(ORACLE_EL[type(el)](), m.coords, m.elements, oracle_density(density))
    fe = None
    if f_ext:
        fe = np.random.default_rng(3).standard_normal(prob.n) * 1e-2
        prob.f_ext = fe
    E = op.energy(density, external_force=fe)
    return op, E, prob


# ------------------------------------------------------------------ KATs ----

def test_operator_kats_on_gpu():
    k = GOLDEN["operator_two_triangles"]
    op = ta.Operator(ta.Mesh(np.array(k["nodes"]), np.array(k["elements"])), ta.Tri3())
    np.testing.assert_allclose(to_np(op.eval(np.array(k["eval_input"]))), k["eval_expected"],
                               rtol=RTOL)
    X = np.array(k["nodes"])
    np.testing.assert_allclose(to_np(op.grad(X @ np.array(k["grad_coeffs"]))), k["grad_expected"],
                               rtol=RTOL)
    np.testing.assert_allclose(to_np(op.integrate_per_element(np.ones(4))), k["areas_expected"],
                               rtol=RTOL)
    np.testing.assert_allclose(float(op.integrate(np.ones(4))), 1.0, rtol=RTOL)
    np.testing.assert_allclose(float(op.integrate(1.0)), 1.0, rtol=RTOL)
    q = np.full((2, 1), k["quad_value"])
    np.testing.assert_allclose(to_np(op.integrate_per_element(q)), k["quad_integral_expected"],
                               rtol=RTOL)
    np.testing.assert_allclose(to_np(op.get_integration_weights()), [[0.5], [0.5]], rtol=RTOL)


@pytest.mark.parametrize("cls", [ta.Tri3, ta.Tetrahedron4, ta.Hexahedron8])
def test_element_linear_fields_exact_on_gpu(cls):
    """tests/test_element.py:45-147 through Operator.grad on a one-element mesh."""
    el = cls()
    X = el._reference_nodes()
    op = ta.Operator(ta.Mesh(X, np.arange(el.nne)[None, :]), el)
    d = el.dim
    rng = np.random.default_rng(0)
    a = np.array([2.0, 3.0, 4.0][:d])
    g = to_np(op.grad(X @ a))[0]
    np.testing.assert_allclose(g, np.tile(a, (g.shape[0], 1)), atol=1e-12)
    A = rng.standard_normal((d, d))
    g = to_np(op.grad(X @ A.T))[0]
    for q in range(g.shape[0]):
        np.testing.assert_allclose(g[q], A, atol=1e-12)
    B = rng.standard_normal((2, 2, d))
    g = to_np(op.grad(np.einsum("ijk,nk->nij", B, X)))[0]
    for q in range(g.shape[0]):
        np.testing.assert_allclose(g[q], B, atol=1e-12)


def test_lifter_kats_on_gpu():
    k = GOLDEN["lifted_fixed_sum"]
    L = lifter.Lifter(k["size"], lifter.Fixed(k["fixed"], 0.0))
    u = L.lift_from_zeros(np.array(k["u_reduced"]))
    np.testing.assert_array_equal(to_np(u), k["lift_expected"])
    assert float(u.sum()) == k["sum_expected"]
    np.testing.assert_array_equal(to_np(L.reduce(u)), k["u_reduced"])
    k = GOLDEN["reduce_adjoint_fixed"]
    L = lifter.Lifter(k["size"], lifter.Fixed(k["fixed"], 0.0))
    np.testing.assert_array_equal(to_np(L.reduce_adjoint(np.array(k["r_full"]))), k["expected"])
    L = lifter.Lifter(4, lifter.Fixed([0, 3], lifter.RuntimeValue("top")))
    u = L.at["top"].set(np.array([1.5, -2.0])).lift(np.array([7.0, 8.0]), np.zeros(4))
    np.testing.assert_array_equal(to_np(u), [1.5, 7.0, 8.0, -2.0])
    dbl = lifter.lifted(lambda x: x * 2, output="primal")
    np.testing.assert_array_equal(to_np(dbl(L.at["top"].set(0.0), np.array([10.0, 20.0]))),
                                  [20.0, 40.0])


# ------------------------------------------------------------ operator ------

@pytest.mark.parametrize("kind", ["tri3", "tet4", "hex8"])
def test_eval_grad_weights_integrate(kind):
    m, el = make_mesh(kind, 4)
    op = ta.Operator(m, el)
    oe = ORACLE_EL[type(el)]()
    rng = np.random.default_rng(1)
    u = rng.standard_normal((m.n_nodes, 3))
    assert rel_err(to_np(op.get_integration_weights()),
                   fem.integration_weights(oe, m.coords, m.elements)) < RTOL
    assert rel_err(to_np(op.eval(u)), fem.op_eval(oe, m.coords, m.elements, u)) < RTOL
    assert rel_err(to_np(op.grad(u)), fem.op_grad(oe, m.coords, m.elements, u)) < RTOL
    assert rel_err(to_np(op.grad(u[:, 0])), fem.op_grad(oe, m.coords, m.elements, u[:, 0])) < RTOL
    assert rel_err(to_np(op.integrate_per_element(u)),
                   fem.integrate_per_element(oe, m.coords, m.elements, u)) < RTOL
    assert rel_err(to_np(op.integrate(u)), fem.integrate(oe, m.coords, m.elements, u)) < RTOL
    qv = rng.standard_normal((m.n_elements, len(el.quad_weights), 2))
    assert rel_err(to_np(op.integrate(qv)), fem.integrate(oe, m.coords, m.elements, qv)) < RTOL
    assert rel_err(to_np(op.quads()), fem.op_eval(oe, m.coords, m.elements, m.coords)) < RTOL


DENSITIES = [energies.Poisson(1.0), energies.LinearElasticity(0.38, 0.58),
             energies.NeoHookean(1.0, 10.0)]
CASES = [(k, d) for k in ("tri3", "tet4", "hex8", "hex8-affine") for d in DENSITIES]


@pytest.mark.parametrize("kind,density", CASES, ids=[f"{k}-{type(d).__name__}" for k, d in CASES])
def test_energy_residual_hvp_parity(kind, density):
    op, E, prob = pair(kind, density, f_ext=True)
    rng = np.random.default_rng(2)
    # Neo-Hookean needs det F > 0 in every element: keep the state small
    amp = 0.01 if isinstance(density, energies.NeoHookean) else 0.05
    u = amp * rng.standard_normal(prob.n)
    v = rng.standard_normal(prob.n)
    assert np.isfinite(prob.energy(u))
    e = float(E(u))
    # 1e-12 relative to the magnitude of the summed terms (sum |psi w| + sum |f_i u_i|):
    # the energy itself can cancel (the external work has either sign)
    G = prob.grad_q(u)
    uq = np.einsum("qn,enc->eqc", prob.Nq, prob._ue(u))
    scale = np.sum(np.abs(prob.density.psi(G, uq) * prob.w)) + np.sum(np.abs(prob.f_ext * u))
    assert abs(e - prob.energy(u)) <= RTOL * scale
    assert rel_err(to_np(E.grad(u)), prob.residual(u)) < RTOL
    assert rel_err(to_np(E.hvp(u, v)), prob.hvp(u, v)) < RTOL
    assert rel_err(to_np(E.hessian_diag(u)), prob.hessian_diag(u)) < RTOL
    r, y = ta.jvp(ta.grad(E), (u,), (v,))
    assert rel_err(to_np(y), prob.hvp(u, v)) < RTOL
    # FP32 Hessian-vector product: same operator in single precision
    t = __import__("torch")
    u32 = t.tensor(u, dtype=t.float32, device="cuda")
    v32 = t.tensor(v, dtype=t.float32, device="cuda")
    assert rel_err(to_np(E.hvp(u32, v32)), prob.hvp(u, v)) < 2e-5
    # psi at quadrature points through the density callable
    psi = density(op.grad(u.reshape(-1, prob.nc) if prob.nc > 1 else u),
                  values=op.eval(u) if density.needs_values else None)
    w = fem.integration_weights(prob.elem, prob.coords, prob.elements)
    prob.f_ext = None
    scale = np.sum(np.abs(prob.density.psi(G, uq) * prob.w))
    assert abs(float((psi * op.get_integration_weights()).sum()) - prob.energy(u)) \
        <= RTOL * scale
    assert w.shape == tuple(psi.shape[:2])


@pytest.mark.parametrize("kind,n", [("tri3", 16), ("tet4", 8), ("hex8", 9), ("hex8-affine", 9)])
def test_assembly_over_many_tiles(kind, n):
    """Residual, HVP and diagonal on meshes of several 256-element tiles (nodes
    shared across tiles are summed by the slot gather) against the oracle."""
    density = energies.NeoHookean(1.0, 10.0)
    op, E, prob = pair(kind, density, n=n, f_ext=True)
    assert op.mesh.elements.shape[0] > 2 * 256
    rng = np.random.default_rng(21)
    # Neo-Hookean needs det F > 0: the perturbation scales with the cell size
    u = 0.04 / n * rng.standard_normal(prob.n)
    v = rng.standard_normal(prob.n)
    assert np.isfinite(prob.residual(u)).all()
    assert rel_err(to_np(E.grad(u)), prob.residual(u)) < RTOL
    assert rel_err(to_np(E.hvp(u, v)), prob.hvp(u, v)) < RTOL
    assert rel_err(to_np(E.hessian_diag(u)), prob.hessian_diag(u)) < RTOL
    t = __import__("torch")
    u32 = t.tensor(u, dtype=t.float32, device="cuda")
    v32 = t.tensor(v, dtype=t.float32, device="cuda")
    assert rel_err(to_np(E.hvp(u32, v32)), prob.hvp(u, v)) < 2e-5
    # integrals: per element and total only (no per-element buffer)
    vals = rng.standard_normal((op.mesh.coords.shape[0], 5))
    pe = to_np(op.integrate_per_element(vals))
    tot = to_np(op.integrate(vals))
    ref = fem.integrate_per_element(prob.elem, prob.coords, prob.elements, vals)
    assert rel_err(pe, ref) < RTOL
    assert rel_err(tot, ref.sum(axis=0)) < RTOL


def test_residual_and_hessian_are_deterministic():
    op, E, prob = pair("hex8", energies.NeoHookean(1.0, 10.0), n=6)
    u = 0.01 * np.random.default_rng(4).standard_normal(prob.n)
    r1, r2 = to_np(E.grad(u)), to_np(E.grad(u))
    assert np.isfinite(r1).all()
    assert np.array_equal(r1, r2)
    S = sparse.pattern_from_mesh(op.mesh, 3)
    cm = sparse.ColoredMatrix.from_csr(S)
    J = sparse.jacfwd(E.grad, cm)
    assert np.array_equal(to_np(J(u).data), to_np(J(u).data))


# ------------------------------------------------------ colored Hessian -----

def test_colored_hessian_tri3_kat():
    """tests/test_sparse.py:49-80: colored == dense Hessian; linearized primal == grad."""
    m = ta.Mesh.unit_square(8, 8)
    op = ta.Operator(m, ta.Tri3())
    E = op.energy(energies.LinearElasticity(1.0, 0.0))
    S = sparse.pattern_from_mesh(m, 2)
    cm = sparse.ColoredMatrix.from_csr(S)
    u0 = np.zeros(m.n_nodes * 2)
    K = sparse.jacfwd(ta.grad(E), cm, color_batch_size=10)(u0)
    prim, K2 = sparse.linearized_jacfwd(ta.grad(E), cm, color_batch_size=10)(u0)
    prob = fem.Problem(fem.Tri3(), m.coords, m.elements, fem.LinearElasticity(1.0, 0.0))
    Hd = np.stack([prob.hvp(u0, np.eye(prob.n)[j]) for j in range(prob.n)], axis=1)
    np.testing.assert_allclose(K.to_dense(), Hd, rtol=1e-7, atol=1e-12)
    np.testing.assert_allclose(K2.to_dense(), Hd, rtol=1e-7, atol=1e-12)
    np.testing.assert_allclose(to_np(prim), prob.residual(u0), atol=1e-15)
    rows, cc = sparse.compute_rows_cols(cm)
    ro, cco = fem.compute_rows_cols(S.indptr, S.indices, to_np(cm.colors))
    np.testing.assert_array_equal(to_np(rows), ro)
    np.testing.assert_array_equal(to_np(cc), cco)


HESS_CASES = [("tri3", energies.Poisson(1.0)), ("tri3", energies.NeoHookean(1.0, 2.0)),
              ("tet4", energies.NeoHookean(1.0, 10.0)), ("tet4", energies.LinearElasticity(0.4, 0.6)),
              ("hex8", energies.LinearElasticity(0.38, 0.58)), ("hex8", energies.NeoHookean(1.0, 10.0))]


@pytest.mark.parametrize("kind,density", HESS_CASES,
                         ids=[f"{k}-{type(d).__name__}" for k, d in HESS_CASES])
def test_colored_hessian_with_lifter_parity(kind, density):
    op, E, prob = pair(kind, density, n=4)
    nc = prob.nc
    m = op.mesh
    x0 = np.where(np.isclose(m.coords[:, 0], 0.0))[0]
    x1 = np.where(np.isclose(m.coords[:, 0], 1.0))[0]
    fixed = np.concatenate([(x0[:, None] * nc + np.arange(nc)).ravel(), x1 * nc])
    L = lifter.Lifter(prob.n, lifter.Fixed(fixed, lifter.RuntimeValue("g", default=0.0)))
    L = L.at["g"].set(0.01)
    Lo = fem.Lifter(prob.n, fem.Fixed(fixed, 0.01))
    S = L.adapt_sparsity(sparse.pattern_from_mesh(m, nc))
    So = Lo.adapt_sparsity(fem.pattern_from_mesh(m.elements, m.n_nodes, nc))
    So.sort_indices()
    cm = sparse.ColoredMatrix.from_csr(S)
    colors_o = fem.distance2_colors(So.indptr, So.indices, So.shape[0])
    np.testing.assert_array_equal(to_np(cm.colors), colors_o)          # bit-exact
    np.testing.assert_array_equal(to_np(cm.indptr), So.indptr)
    np.testing.assert_array_equal(to_np(cm.indices), So.indices)
    rng = np.random.default_rng(6)
    ured = 0.03 * rng.standard_normal(L.size_reduced)
    Ered = E.reduced(L)
    K = sparse.jacfwd(Ered.grad, cm)(ured)
    uf = Lo.lift_from_zeros(ured)
    Ko = fem.hessian_csr(prob, Lo, uf, So, colors_o)   # literal one-JVP-per-color oracle
    np.testing.assert_array_equal(to_np(K.indices), Ko.indices)
    assert rel_err(to_np(K.data), Ko.data) < RTOL
    prim, K2 = sparse.linearized_jacfwd(Ered.grad, cm)(ured)
    assert rel_err(to_np(prim), Lo.reduce_adjoint(prob.residual(uf))) < RTOL
    assert np.array_equal(to_np(K2.data), to_np(K.data))
    # SpMV on the assembled matrix equals the matrix-free HVP
    v = rng.standard_normal(L.size_reduced)
    assert rel_err(to_np(K @ v), to_np(Ered.hvp(ured, v))) < 1e-11


# ------------------------------------------------------------ solvers -------

def test_pcg_matches_spsolve_and_oracle_iterations():
    m = ta.Mesh.unit_cube(6)
    op = ta.Operator(m, ta.Hexahedron8())
    E = op.energy(energies.LinearElasticity(0.4, 0.6))
    x0n = np.where(np.isclose(m.coords[:, 0], 0.0))[0]
    L = lifter.Lifter(m.n_nodes * 3, lifter.Fixed((x0n[:, None] * 3 + np.arange(3)).ravel()))
    cm = sparse.ColoredMatrix.from_csr(L.adapt_sparsity(sparse.pattern_from_mesh(m, 3)))
    K = sparse.jacfwd(E.reduced(L).grad, cm)(np.zeros(L.size_reduced))
    b = np.random.default_rng(5).standard_normal(L.size_reduced)
    x, info = solvers.pcg(K, b, tol=1e-10, chunk=8)
    assert info.converged and info.rel_residual <= 1e-10
    Kh = K.to_csr()
    xo, ito, relo = fem.pcg(Kh, b, tol=1e-10)
    assert abs(info.iterations - ito) <= 2
    # the GPU solver runs the single-reduction form: same recurrence as the oracle's
    x1, it1, rel1 = fem.pcg_single_reduction(Kh, b, tol=1e-10)
    assert abs(info.iterations - it1) <= 1
    xs = spla.spsolve(Kh.tocsc(), b)
    assert rel_err(to_np(x), xs) < 1e-8
    true_rel = np.linalg.norm(b - Kh @ to_np(x)) / np.linalg.norm(b)
    assert true_rel < 1e-9
    # the node-block view gives the same operator and the same solve
    view = K.block_view(3)
    assert view is not None and view[0].numel() == K.nnz // 9
    v = np.random.default_rng(9).standard_normal(L.size_reduced)
    assert rel_err(to_np(K.matvec(v, block_size=3)), Kh @ v) < 1e-13
    assert rel_err(to_np(K.matvec(v)), Kh @ v) < 1e-13
    xb, infob = solvers.pcg(K, b, tol=1e-10, chunk=8, block_size=3)
    assert infob.converged and abs(infob.iterations - info.iterations) <= 1
    assert rel_err(to_np(xb), to_np(x)) < 1e-9
    # a pattern that is not node-blocked (one x-component constrained) falls back to CSR
    L2 = lifter.Lifter(m.n_nodes * 3, lifter.Fixed(np.concatenate(
        [(x0n[:, None] * 3 + np.arange(3)).ravel(), [3 * (m.n_nodes - 1)]])))
    cm2 = sparse.ColoredMatrix.from_csr(L2.adapt_sparsity(sparse.pattern_from_mesh(m, 3)))
    assert cm2.block_view(3) is None
    # zero right-hand side and maxiter limit
    x, info = solvers.pcg(K, np.zeros_like(b))
    assert info.converged and info.iterations == 0 and float(x.abs().max()) == 0.0
    x, info = solvers.pcg(K, b, maxiter=3, chunk=2)
    assert info.iterations == 3 and not info.converged


@pytest.mark.parametrize("cells,n_expected", [(24, 45000), (56, 545832)])
def test_pcg_graph_path_both_layouts(cells, n_expected):
    """Systems above the one-workgroup size run the graph-captured PCG.  At
    n = 45,000 both operators' block shares fit the CG step's re-sum; at
    n = 545,832 they exceed it and pass through the 64-block share sum.  The
    node-block view and the sliced ELL copy solve the same system; a pattern
    with one dof of the x = 1 face constrained (not node-blocked) solves
    through the sliced ELL."""
    import torch

    m = ta.Mesh.unit_cube(cells)
    op = ta.Operator(m, ta.Hexahedron8())
    E = op.energy(energies.LinearElasticity(0.4, 0.6))
    x0n = np.where(np.isclose(m.coords[:, 0], 0.0))[0]
    x1n = np.where(np.isclose(m.coords[:, 0], 1.0))[0]
    L = lifter.Lifter(m.n_nodes * 3, lifter.Fixed((x0n[:, None] * 3 + np.arange(3)).ravel()))
    cm = sparse.ColoredMatrix.from_csr(L.adapt_sparsity(sparse.pattern_from_mesh(m, 3)))
    K = sparse.jacfwd(E.reduced(L).grad, cm)(np.zeros(L.size_reduced))
    n = L.size_reduced
    assert n == n_expected and n % 64 != 0
    b = np.random.default_rng(11).standard_normal(n)
    bt = torch.as_tensor(b, device="cuda")
    xs, infos = solvers.pcg(K, b, tol=1e-10, chunk=16)
    xb, infob = solvers.pcg(K, b, tol=1e-10, chunk=16, block_size=3)
    assert infos.converged and infob.converged
    assert abs(infos.iterations - infob.iterations) <= 1
    for x in (xs, xb):
        res = float(torch.linalg.norm(bt - K.matvec(x)) / torch.linalg.norm(bt))
        assert res <= 2e-10
    assert rel_err(to_np(xs), to_np(xb)) < 1e-8
    L2 = lifter.Lifter(m.n_nodes * 3, lifter.Fixed(np.concatenate(
        [(x0n[:, None] * 3 + np.arange(3)).ravel(), x1n * 3])))
    cm2 = sparse.ColoredMatrix.from_csr(L2.adapt_sparsity(sparse.pattern_from_mesh(m, 3)))
    assert cm2.block_view(3) is None
    K2 = sparse.jacfwd(E.reduced(L2).grad, cm2)(np.zeros(L2.size_reduced))
    b2 = torch.as_tensor(np.random.default_rng(12).standard_normal(L2.size_reduced), device="cuda")
    x2, info2 = solvers.pcg(K2, b2, tol=1e-10, chunk=16)
    assert info2.converged
    assert float(torch.linalg.norm(b2 - K2.matvec(x2)) / torch.linalg.norm(b2)) <= 2e-10


@pytest.mark.parametrize("n", [16, 64], ids=["16x16", "c1-64x64"])
def test_newton_poisson_tri3_matches_oracle(n):
    """C1 (Tri3 Poisson, f = 1, u = 0 on the boundary) at 16^2 and at its full
    64^2 size: one Newton step (quadratic energy), solution vs the oracle's."""
    m = ta.Mesh.unit_square(n, n)
    op = ta.Operator(m, ta.Tri3())
    E = op.energy(energies.Poisson(1.0))
    bnd = np.where(np.isclose(m.coords, 0).any(1) | np.isclose(m.coords, 1).any(1))[0]
    L = lifter.Lifter(m.n_nodes, lifter.Fixed(bnd, 0.0))
    solver = solvers.Newton(E, L, mode="csr", tol=1e-10, lin_tol=1e-12)
    t = __import__("torch")
    u = t.zeros(m.n_nodes, dtype=t.float64, device="cuda")
    st = solver.solve(u)
    assert st.converged and st.newton_iterations == 1
    prob = fem.Problem(fem.Tri3(), m.coords, m.elements, fem.Poisson(1.0))
    Lo = fem.Lifter(m.n_nodes, fem.Fixed(bnd))
    So = Lo.adapt_sparsity(fem.pattern_from_mesh(m.elements, m.n_nodes, 1))
    co = fem.distance2_colors(So.indptr, So.indices, So.shape[0])
    uo, its, cg, hist = fem.newton(prob, Lo, np.zeros(m.n_nodes), So, co, tol=1e-12, lin_tol=1e-12)
    assert rel_err(to_np(u), uo) < 1e-10


def test_newton_mixed_precision_neohookean():
    """FP32 matrix-free CG inside FP64 Newton on a Hex8 Neo-Hookean cube, two load
    steps set through lifter.at[key].set: the oracle's FP64 residual at the GPU
    state is <= 1e-8 of the oracle residual at the step's start, and the state
    is the oracle Newton solution of the same load to 1e-6 relative (both stop
    at a relative residual; their states differ by the conditioning times it)."""
    t = __import__("torch")
    m = ta.Mesh.unit_cube(5)
    op = ta.Operator(m, ta.Hexahedron8())
    E = op.energy(energies.NeoHookean(1.0, 10.0))
    x0n = np.where(np.isclose(m.coords[:, 0], 0.0))[0]
    x1n = np.where(np.isclose(m.coords[:, 0], 1.0))[0]
    fixed = np.concatenate([(x0n[:, None] * 3 + np.arange(3)).ravel(), x1n * 3])
    L = lifter.Lifter(m.n_nodes * 3, lifter.Fixed(fixed, lifter.RuntimeValue("s")))
    vals = lambda s: np.concatenate([np.zeros(x0n.size * 3), np.full(x1n.size, s)])  # noqa: E731
    prob = fem.Problem(fem.Hex8(), m.coords, m.elements, fem.NeoHookean(1.0, 10.0))
    So = fem.Lifter(m.n_nodes * 3, fem.Fixed(fixed)).adapt_sparsity(
        fem.pattern_from_mesh(m.elements, m.n_nodes, 3))
    co = fem.distance2_colors(So.indptr, So.indices, So.shape[0])
    um = t.zeros(m.n_nodes * 3, dtype=t.float64, device="cuda")
    mixed = solvers.Newton(E, L.at["s"].set(vals(0.0)), mode="mixed", tol=1e-8, inner_tol=1e-3)
    for s in (0.05, 0.1):
        Lo = fem.Lifter(m.n_nodes * 3, fem.Fixed(fixed, vals(s)))
        start = to_np(um)
        start = Lo.lift(Lo.reduce(start), start)
        st_m = mixed.solve(um, L.at["s"].set(vals(s)))
        assert st_m.converged
        assert st_m.residual_history[-1] <= 1e-8 * st_m.residual_history[0]
        uo, _, _, hist = fem.newton(prob, Lo, start, So, co, tol=1e-11, lin_tol=1e-12)
        assert hist[-1] <= 1e-10 * hist[0]
        ug = to_np(um)
        r0 = np.linalg.norm(Lo.reduce_adjoint(prob.residual(start)))
        assert np.linalg.norm(Lo.reduce_adjoint(prob.residual(ug))) <= 1e-8 * r0
        assert rel_err(ug, uo) < 1e-6


# ------------------------------------------------------------- periodic ----

def _periodic_case(nx, ny, density):
    """Tri3 square, Fixed bottom (u = 0) and top (u_y = 0.01), left edge periodic
    with the right edge (tests/test_sparse_tracer.py:162-216 layout)."""
    m = ta.Mesh.unit_square(nx, ny)
    X = m.coords
    bot = np.where(X[:, 1] == 0.0)[0]
    top = np.where(X[:, 1] == 1.0)[0]
    left = np.where(X[:, 0] == 0.0)[0]
    left = left[np.argsort(X[left, 1])]
    right = np.where(X[:, 0] == 1.0)[0]
    right = right[np.argsort(X[right, 1])]
    c = np.arange(2)
    fixed = (bot[:, None] * 2 + c).ravel()
    slaves, masters = (left[:, None] * 2 + c).ravel(), (right[:, None] * 2 + c).ravel()
    L = lifter.Lifter(m.n_nodes * 2, lifter.Fixed(fixed, 0.0), lifter.Fixed(top * 2 + 1, 0.01),
                      lifter.Periodic(slaves, masters))
    Lo = fem.Lifter(m.n_nodes * 2, fem.Fixed(fixed, 0.0), fem.Fixed(top * 2 + 1, 0.01),
                    fem.Periodic(slaves, masters))
    op = ta.Operator(m, ta.Tri3())
    E = op.energy(density)
    prob = fem.Problem(fem.Tri3(), X, m.elements, oracle_density(density))
    return m, L, Lo, E, prob, slaves, masters


def test_periodic_lifter_kats_on_gpu():
    """tests/test_lifter.py:26-41, tests/test_lifter_residual.py:7-78."""
    k = GOLDEN["lifter_fixed_periodic"]
    L = lifter.Lifter(k["size"], lifter.Fixed(k["fixed"]),
                      lifter.Periodic(k["periodic_dofs"], k["periodic_master"]))
    u = L.lift_from_zeros(np.array(k["u_reduced"]))
    np.testing.assert_array_equal(to_np(u), k["lift_expected"])
    np.testing.assert_array_equal(to_np(L.reduce(u)), k["u_reduced"])
    for name in ("reduce_adjoint_periodic", "reduce_adjoint_periodic_fixed"):
        k = GOLDEN[name]
        cons = [lifter.Periodic(k["periodic_dofs"], k["periodic_master"])]
        if "fixed" in k:
            cons.append(lifter.Fixed(k["fixed"]))
        L = lifter.Lifter(k["size"], *cons)
        np.testing.assert_array_equal(to_np(L.reduce_adjoint(np.array(k["r_full"]))),
                                      k["expected"])
    # tests/test_lifter_residual.py:39-56: lifted dual output
    L = lifter.Lifter(4, lifter.Periodic([2], [1]))
    f = lifter.lifted(lambda u: u * 2.0, output="dual")
    np.testing.assert_array_equal(to_np(f(L, np.array([1.0, 2.0, 3.0]))), [2.0, 8.0, 6.0])


@pytest.mark.parametrize("density", [energies.LinearElasticity(0.4, 0.6),
                                     energies.NeoHookean(1.0, 1.0)],
                         ids=["LinearElasticity", "NeoHookean"])
def test_periodic_residual_hvp_hessian_newton(density):
    m, L, Lo, E, prob, slaves, masters = _periodic_case(6, 5, density)
    red = E.reduced(L)
    rng = np.random.default_rng(7)
    ur = 0.01 * rng.standard_normal(L.size_reduced)
    vr = rng.standard_normal(L.size_reduced)
    uf = Lo.lift_from_zeros(ur)
    np.testing.assert_allclose(to_np(red.lift(ur)), uf, rtol=0, atol=1e-15)
    r_or = Lo.reduce_adjoint(prob.residual(uf))
    assert rel_err(to_np(red.grad(ur)), r_or) < RTOL
    y_or = Lo.reduce_adjoint(prob.hvp(uf, Lo.lift_linear(vr)))
    assert rel_err(to_np(red.hvp(ur, vr)), y_or) < RTOL
    # colored reduced Hessian on the augmented + reduced pattern
    S = L.adapt_sparsity(sparse.pattern_from_mesh(m, 2))
    cm = sparse.ColoredMatrix.from_csr(S)
    K = sparse.jacfwd(red.grad, cm)(ur)
    So = Lo.adapt_sparsity(fem.pattern_from_mesh(m.elements, m.n_nodes, 2))
    np.testing.assert_array_equal(to_np(cm.indices), So.indices)
    Ko = fem.hessian_csr(prob, Lo, uf, So, to_np(cm.colors))
    assert rel_err(to_np(K.data), Ko.data) < RTOL
    # Newton (colored CSR + PCG) against the oracle's Newton
    u = t_zeros(m.n_nodes * 2)
    st = solvers.Newton(E, L, mode="csr", tol=1e-10, lin_tol=1e-12).solve(u)
    assert st.converged
    uo, *_ = fem.newton(prob, Lo, Lo.lift_from_zeros(np.zeros(Lo.size_reduced)), So,
                        to_np(cm.colors), tol=1e-10, lin_tol=1e-12)
    un = to_np(u)
    np.testing.assert_array_equal(un[slaves], un[masters])
    assert np.abs(un - uo).max() <= 1e-8 * np.abs(uo).max()


@pytest.mark.parametrize("density", [energies.LinearElasticity(0.4, 0.6),
                                     energies.NeoHookean(1.0, 1.0)],
                         ids=["LinearElasticity", "NeoHookean"])
def test_periodic_mixed_newton_matches_oracle(density):
    """Mixed Newton (FP32 matrix-free PCG on P^T H P, FP64 residual) under
    Periodic + Fixed constraints: the FP32 reduced operator applied as a whole
    (tatva_pcg_matfree_f32_mapped) reaches the oracle's Newton solution, and
    every slave equals its master."""
    m, L, Lo, E, prob, slaves, masters = _periodic_case(6, 5, density)
    S = L.adapt_sparsity(sparse.pattern_from_mesh(m, 2))
    cm = sparse.ColoredMatrix.from_csr(S)
    So = Lo.adapt_sparsity(fem.pattern_from_mesh(m.elements, m.n_nodes, 2))
    uo, *_ = fem.newton(prob, Lo, Lo.lift_from_zeros(np.zeros(Lo.size_reduced)), So,
                        to_np(cm.colors), tol=1e-10, lin_tol=1e-12)
    u = t_zeros(m.n_nodes * 2)
    st = solvers.Newton(E, L, mode="mixed", tol=1e-10).solve(u)
    assert st.converged and st.cg_iterations > 0
    un = to_np(u)
    np.testing.assert_array_equal(un[slaves], un[masters])
    assert np.abs(un - uo).max() <= 1e-7 * np.abs(uo).max()


def test_periodic_hessian_when_slave_and_master_share_elements():
    """One cell across: every element holds a slave and its master (the serial
    accumulation path of the mapped colored Hessian)."""
    m, L, Lo, E, prob, slaves, masters = _periodic_case(1, 4, energies.LinearElasticity(0.4, 0.6))
    red = E.reduced(L)
    S = L.adapt_sparsity(sparse.pattern_from_mesh(m, 2))
    cm = sparse.ColoredMatrix.from_csr(S)
    ur = np.zeros(L.size_reduced)
    K = sparse.jacfwd(red.grad, cm)(ur)
    So = Lo.adapt_sparsity(fem.pattern_from_mesh(m.elements, m.n_nodes, 2))
    Ko = fem.hessian_csr(prob, Lo, Lo.lift_from_zeros(ur), So, to_np(cm.colors))
    assert rel_err(to_np(K.data), Ko.data) < RTOL


def t_zeros(n):
    import torch
    return torch.zeros(n, dtype=torch.float64, device="cuda")


# ----------------------------------------------------------- projection ----

def _tri3_quadratic_rule():
    return np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]]), np.full(3, 1 / 6)


def test_project_reproduces_linear_and_fe_fields():
    """tests/test_operator_projection.py:21-86 on Tri3 (3-point rule) and Hex8:
    scalar, vector (scalar pattern and coupled 2-dof pattern), tensor and the
    default pattern; any field in the FE space is reproduced."""
    qp, qw = _tri3_quadratic_rule()
    m = ta.Mesh.unit_square(2, 2)
    op = ta.Operator(m, ta.Tri3(qp, qw))
    q = to_np(op.quads())
    X = m.coords
    s = op.project(q[:, :, 0] + q[:, :, 1], sparse.ColoredMatrix.from_csr(
        sparse.pattern_from_mesh(m, 1)))
    np.testing.assert_allclose(to_np(s), X[:, 0] + X[:, 1], atol=1e-10)
    v = op.project(q)
    assert tuple(v.shape) == (m.n_nodes, 2)
    np.testing.assert_allclose(to_np(v), X, atol=1e-10)
    vc = op.project(q, sparse.ColoredMatrix.from_csr(sparse.pattern_from_mesh(m, 2)))
    np.testing.assert_allclose(to_np(vc), X, atol=1e-10)
    T = np.zeros(q.shape[:2] + (2, 2))
    T[..., 0, 0], T[..., 1, 1] = q[..., 0], q[..., 1]
    Tp = to_np(op.project(T))
    np.testing.assert_allclose(Tp[:, 0, 0], X[:, 0], atol=1e-10)
    np.testing.assert_allclose(Tp[:, 1, 1], X[:, 1], atol=1e-10)
    np.testing.assert_allclose(Tp[:, 0, 1], 0.0, atol=1e-10)
    # Hex8 on a jittered cube: a random FE field and the oracle's projection
    mh, el = make_mesh("hex8", 3)
    oph = ta.Operator(mh, el)
    nodal = np.random.default_rng(8).standard_normal((mh.n_nodes, 3))
    f = fem.op_eval(fem.Hex8(), mh.coords, mh.elements, nodal)
    np.testing.assert_allclose(to_np(oph.project(f)), nodal, atol=1e-9)
    bnd = np.where(np.isclose(mh.coords[:, 0], 0.0))[0]
    L = lifter.Lifter(mh.n_nodes, lifter.Fixed(bnd, 0.25))
    Lo = fem.Lifter(mh.n_nodes, fem.Fixed(bnd, 0.25))
    got = to_np(oph.project(f[..., 0], lifter=L))
    want = fem.project(fem.Hex8(), mh.coords, mh.elements, f[..., 0], lifter=Lo)
    assert rel_err(got, want) < 1e-9


def test_project_under_periodic_constraints():
    """Operator.project with Fixed and Periodic constraints (utils.py:225-246):
    the reference's system E_free^T M P (test functions at the free dofs, the
    trial field lifted: not symmetric, solved by BiCGSTAB) against the oracle's
    dense solve, 1e-9, for a constrained field (reproduced exactly) and for a
    random one with non-zero Fixed values (whose constant part the reference's
    system leaves out); every slave equals its master."""
    qp, qw = _tri3_quadratic_rule()
    m = ta.Mesh.unit_square(6, 5)
    op = ta.Operator(m, ta.Tri3(qp, qw))
    tri = fem.Tri3(qp, qw)
    X = m.coords
    left = np.where(np.isclose(X[:, 0], 0.0))[0]
    left = left[np.argsort(X[left, 1])]
    right = np.where(np.isclose(X[:, 0], 1.0))[0]
    right = right[np.argsort(X[right, 1])]
    bottom = np.where(np.isclose(X[:, 1], 0.0))[0]
    slaves = np.setdiff1d(left, bottom)
    masters = right[np.isin(left, slaves)]
    g = np.cos(2 * np.pi * X[:, 0]) * X[:, 1] + X[:, 1] ** 2
    g[slaves] = g[masters]
    for val in (0.0, 0.3):
        L = lifter.Lifter(m.n_nodes, lifter.Fixed(bottom, val), lifter.Periodic(slaves, masters))
        Lo = fem.Lifter(m.n_nodes, fem.Fixed(bottom, val), fem.Periodic(slaves, masters))
        rng = np.random.default_rng(int(val * 10))
        nodal = g if val == 0.0 else rng.standard_normal(m.n_nodes)
        fq = fem.op_eval(tri, X, m.elements, nodal)
        got = to_np(op.project(fq, lifter=L))
        want = fem.project(tri, X, m.elements, fq, lifter=Lo)
        assert rel_err(got, want) < 1e-9, val
        np.testing.assert_array_equal(got[slaves], got[masters])
        np.testing.assert_array_equal(got[bottom], val)
        if val == 0.0:
            assert rel_err(got, g) < 1e-9


def test_interpolate_matches_oracle():
    """operator.py:399-468 on a jittered Tri3 mesh: interpolated values against
    the oracle (same containing element, same Newton step); linear fields exact;
    a point outside raises (operator.py:453-459)."""
    m, el = make_mesh("tri3", 6)
    op = ta.Operator(m, el)
    rng = np.random.default_rng(10)
    pts = np.vstack([rng.uniform(0, 1, (200, 2)), m.coords[::5], [[0.5, 0.5], [0.0, 0.3]]])
    nodal = rng.standard_normal((m.n_nodes, 3))
    got = to_np(op.interpolate(nodal, pts))
    want = fem.interpolate(fem.Tri3(), m.coords, m.elements, nodal, pts)
    assert np.isfinite(got).all()
    # points on shared edges may be located in either neighbour: the field is
    # continuous there, so values agree to rounding
    assert rel_err(got, want) < 1e-10
    lin = 0.5 + m.coords @ np.array([1.5, -2.0])
    np.testing.assert_allclose(to_np(op.interpolate(lin, pts)), 0.5 + pts @ np.array([1.5, -2.0]),
                               atol=1e-12)
    with pytest.raises(RuntimeError):
        op.interpolate(nodal, np.array([[1.2, 0.5]]))


def test_interpolate_reference_kats():
    """tests/test_operator.py:146-167: the sum of the coordinates interpolated at
    four points of the two-triangle mesh (one on the shared diagonal); a point
    outside raises RuntimeError."""
    k, m = GOLDEN["interpolate_two_triangles"], GOLDEN["operator_two_triangles"]
    X = np.array(m["nodes"])
    op = ta.Operator(ta.Mesh(X, np.array(m["elements"])), ta.Tri3())
    got = to_np(op.interpolate(X.sum(axis=1), np.array(k["points"])))
    np.testing.assert_allclose(got, k["expected"], rtol=1e-12)
    with pytest.raises(RuntimeError, match="outside the mesh"):
        op.interpolate(np.ones(X.shape[0]), np.array(k["outside_points"]))


def test_interpolate_grid_location_on_a_graded_mesh():
    """Point location through the mesh's cell grid on a 3 x 1 Tri3 mesh with
    graded (x -> 3 (x/3)^1.7) and jittered nodes, so elements differ in size
    and many span several cells: random points, every vertex and edge
    midpoints (shared by several triangles) give the oracle's values, found by
    the first-element rule of operator.py:399-468 over all elements."""
    m0 = ta.Mesh.rectangle((0.0, 3.0), (0.0, 1.0), 60, 24)
    X = m0.coords.copy()
    X[:, 0] = 3.0 * (X[:, 0] / 3.0) ** 1.7
    rng = np.random.default_rng(12)
    inner = ~np.isclose(m0.coords[:, 0], 0.0) & ~np.isclose(m0.coords[:, 0], 3.0) & \
        ~np.isclose(m0.coords[:, 1], 0.0) & ~np.isclose(m0.coords[:, 1], 1.0)
    X[:, 1] += inner * rng.uniform(-0.25, 0.25, X.shape[0]) / 24
    m = ta.Mesh(X, m0.elements)
    op = ta.Operator(m, ta.Tri3())
    mids = 0.5 * (X[m.elements[::2, 0]] + X[m.elements[::2, 2]])
    pts = np.vstack([rng.uniform([0.0, 0.0], [3.0, 1.0], (2500, 2)), X, mids])
    nodal = rng.standard_normal((m.n_nodes, 2))
    got = to_np(op.interpolate(nodal, pts))
    want = fem.interpolate(fem.Tri3(), X, m.elements, nodal, pts, batched=True)
    assert np.isfinite(want).all()
    assert rel_err(got, want) < 1e-10
    lin = 0.5 + X @ np.array([1.5, -2.0])
    np.testing.assert_allclose(to_np(op.interpolate(lin, pts)), 0.5 + pts @ np.array([1.5, -2.0]),
                               atol=1e-11)
    with pytest.raises(RuntimeError):
        op.interpolate(nodal, np.vstack([pts[:10], [[3.2, 0.5]], [[1.0, -0.1]]]))
