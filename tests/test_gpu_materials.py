"""Per-element material parameters (the element quantities of tatva's
Operator.map(..., element_quantity=...), tatva/operator.py:229,256) and the
Operator.map / map_over_elements dispatch, on the GPU against the oracle.

Two-material meshes: elements with centroid x < 0.5 have one (mu, lambda), the
others another.  FP64 kernels agree with the oracle to 1e-12 relative, FP32 ones
to 2e-5 (the tolerances of the uniform-parameter parity tests)."""

import numpy as np
import pytest

import tatva_amd as ta
from oracle import fem
from tatva_amd import energies, lifter, solvers, sparse

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def to_np(t):
    return t.detach().cpu().numpy()


def rel_err(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def two_material(kind):
    if kind == "tet4":
        m = ta.Mesh.unit_cube(4, type="tetrahedron", jitter=0.2, seed=3)
        el, oel = ta.Tetrahedron4(), fem.Tet4()
    elif kind == "hex8":
        m = ta.Mesh.unit_cube(4, jitter=0.2, seed=4)
        el, oel = ta.Hexahedron8(), fem.Hex8()
    else:
        m = ta.Mesh.unit_cube(5)
        el, oel = ta.Hexahedron8(), fem.Hex8()
    left = m.coords[m.elements].mean(axis=1)[:, 0] < 0.5
    mu = np.where(left, 1.0, 2.5)
    lam = np.where(left, 10.0, 4.0)
    return m, el, oel, mu, lam


@pytest.mark.parametrize("kind", ["tet4", "hex8", "hex8-affine"])
@pytest.mark.parametrize("dens", ["LinearElasticity", "NeoHookean"])
def test_two_material_energy_residual_hvp_hessian(kind, dens):
    m, el, oel, mu, lam = two_material(kind)
    D = getattr(energies, dens)(mu, lam)
    op = ta.Operator(m, el)
    E = op.energy(D)
    prob = fem.Problem(oel, m.coords, m.elements, getattr(fem, dens)(mu, lam))
    rng = np.random.default_rng(6)
    u = 0.01 * rng.standard_normal(prob.n)
    v = rng.standard_normal(prob.n)
    assert abs(float(E(u)) - prob.energy(u)) <= RTOL * abs(prob.energy(u))
    assert rel_err(to_np(E.grad(u)), prob.residual(u)) < RTOL
    y = prob.hvp(u, v)
    assert rel_err(to_np(E.hvp(u, v)), y) < RTOL
    import torch
    y32 = E.hvp(torch.as_tensor(u, dtype=torch.float32, device="cuda"),
                torch.as_tensor(v, dtype=torch.float32, device="cuda"))
    assert rel_err(to_np(y32), y) < 2e-5
    assert rel_err(to_np(E.hessian_diag(u)), prob.hessian_diag(u)) < RTOL
    # the colored Hessian (one JVP per color) with x = 0 clamped
    x0 = np.where(np.isclose(m.coords[:, 0], 0.0))[0]
    fixed = (x0[:, None] * 3 + np.arange(3)).ravel()
    L = lifter.Lifter(m.n_nodes * 3, lifter.Fixed(fixed, 0.0))
    Lo = fem.Lifter(m.n_nodes * 3, fem.Fixed(fixed))
    cm = sparse.ColoredMatrix.from_csr(L.adapt_sparsity(sparse.pattern_from_mesh(m, 3)))
    uf = Lo.lift_from_zeros(Lo.reduce(u))
    K = sparse.jacfwd(E.reduced(L).grad, cm)(Lo.reduce(u))
    So = Lo.adapt_sparsity(fem.pattern_from_mesh(m.elements, m.n_nodes, 3))
    Ko = fem.hessian_csr(prob, Lo, uf, So, to_np(cm.colors))
    assert rel_err(to_np(K.data), Ko.data) < RTOL
    # the per-element parameters matter: a uniform material gives another residual
    E1 = op.energy(getattr(energies, dens)(1.0, 10.0))
    assert rel_err(to_np(E1.grad(u)), prob.residual(u)) > 1e-3


@pytest.mark.parametrize("mode", ["mixed", "csr"])
def test_two_material_newton_matches_oracle(mode):
    """One stretch load step (u_x = 0.05 on x = 1, x = 0 clamped) of a Neo-Hookean
    two-material Hex8 cube: the GPU Newton state against the oracle's Newton,
    1e-6 relative, and the oracle residual at it <= 1e-8 of the start's."""
    import torch

    m, el, oel, mu, lam = two_material("hex8")
    op = ta.Operator(m, el)
    E = op.energy(energies.NeoHookean(mu, lam))
    x0 = np.where(np.isclose(m.coords[:, 0], 0.0))[0]
    x1 = np.where(np.isclose(m.coords[:, 0], 1.0))[0]
    dofs = np.concatenate([(x0[:, None] * 3 + np.arange(3)).ravel(), x1 * 3])
    vals = np.concatenate([np.zeros(x0.size * 3), np.full(x1.size, 0.05)])
    L = lifter.Lifter(m.n_nodes * 3, lifter.Fixed(dofs, vals))
    Lo = fem.Lifter(m.n_nodes * 3, fem.Fixed(dofs, vals))
    prob = fem.Problem(oel, m.coords, m.elements, fem.NeoHookean(mu, lam))
    u = torch.zeros(m.n_nodes * 3, dtype=torch.float64, device="cuda")
    u[0::3] = torch.as_tensor(0.05 * m.coords[:, 0], device="cuda")
    start = Lo.lift(Lo.reduce(to_np(u)), to_np(u))
    kw = dict(lin_tol=1e-10) if mode == "csr" else dict(inner_tol=1e-3)
    st = solvers.Newton(E, L, mode=mode, tol=1e-8, chunk=16, **kw).solve(u)
    assert st.converged
    So = Lo.adapt_sparsity(fem.pattern_from_mesh(m.elements, m.n_nodes, 3))
    co = fem.distance2_colors(So.indptr, So.indices, So.shape[0])
    uo, _, _, hist = fem.newton(prob, Lo, start, So, co, tol=1e-11, lin_tol=1e-12)
    assert hist[-1] <= 1e-10 * hist[0]
    ug = to_np(u)
    r0 = np.linalg.norm(Lo.reduce_adjoint(prob.residual(start)))
    assert np.linalg.norm(Lo.reduce_adjoint(prob.residual(ug))) <= 1e-8 * r0
    assert rel_err(ug, uo) < 1e-6


def test_operator_map_dispatch():
    """Operator.map / map_over_elements on the GPU: the built-in element
    functions, a density instance, a density class with its parameters as
    element quantities (tatva/operator.py:229,253-258: those arguments are
    passed per element, the others gathered from nodes), per-element
    integration, and the errors for what cannot run on the GPU."""
    m, el, oel, mu, lam = two_material("tet4")
    op = ta.Operator(m, el)
    rng = np.random.default_rng(9)
    u = 0.01 * rng.standard_normal((m.n_nodes, 3))
    G = fem.op_grad(oel, m.coords, m.elements, u)
    assert rel_err(to_np(op.map("grad")(u)), G) < RTOL
    assert rel_err(to_np(op.map("eval")(u)), fem.op_eval(oel, m.coords, m.elements, u)) < RTOL
    w = fem.integration_weights(oel, m.coords, m.elements) / oel.quad_weights[None, :]
    assert rel_err(to_np(op.map("det_jacobian")()), w) < RTOL
    psi_u = fem.NeoHookean(1.0, 10.0).psi(G)
    assert rel_err(to_np(op.map(energies.NeoHookean(1.0, 10.0))(u)), psi_u) < RTOL
    psi_e = fem.NeoHookean(mu, lam).psi(G)
    got = op.map(energies.NeoHookean, element_quantity=(1, 2))(u, mu, lam)
    assert rel_err(to_np(got), psi_e) < RTOL
    # the element quantities may come first: argument 0 and 1 per element, 2 nodal
    got2 = op.map(energies.NeoHookean, element_quantity=(0, 1))(mu, lam, u)
    assert rel_err(to_np(got2), psi_e) < RTOL
    # a scalar field with a per-element source f
    m2 = ta.Mesh.unit_square(6, 5)
    op2 = ta.Operator(m2, ta.Tri3())
    f_e = np.where(m2.coords[m2.elements].mean(axis=1)[:, 0] < 0.5, 1.0, -2.0)
    us = rng.standard_normal(m2.n_nodes)
    Gs = fem.op_grad(fem.Tri3(), m2.coords, m2.elements, us)
    uq = fem.op_eval(fem.Tri3(), m2.coords, m2.elements, us)[..., None]
    want = fem.Poisson(f_e).psi(Gs[:, :, None, :], uq)
    got3 = op2.map(energies.Poisson, element_quantity=(1,))(us, f_e)
    assert rel_err(to_np(got3), want) < RTOL
    # per-element integration of quadrature values
    q = rng.standard_normal((m.n_elements, op.n_quad))
    want_pe = fem.integrate_per_element(oel, m.coords, m.elements, q)
    assert rel_err(to_np(op.map_over_elements("integrate")(q)), want_pe) < RTOL
    assert rel_err(to_np(op.map_over_elements("integrate", element_quantity=(0,))(q)),
                   want_pe) < RTOL
    with pytest.raises(NotImplementedError):
        op.map(lambda xi, ue: ue)
    with pytest.raises(NotImplementedError):
        op.map("grad", element_quantity=(1,))
    with pytest.raises(ValueError, match="element quantities"):
        op.map(energies.NeoHookean)(u)
    with pytest.raises(ValueError, match=f"{m.n_elements} values"):
        op.map(energies.NeoHookean, element_quantity=(1, 2))(u, mu[:-1], lam)
    with pytest.raises(ValueError, match="per-element parameters"):
        op.energy(energies.NeoHookean(mu[:-1], lam[:-1]))
