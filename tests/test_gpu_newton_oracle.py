"""GPU parity of the solver path the mixed-vs-FP64 measurements run: load-stepped
compressible Neo-Hookean Newton, in both modes, against the FP64 oracle.

* ``mode="mixed"``: FP32 matrix-free Jacobi-PCG (tatva_pcg_matfree_f32, Fixed
  constraints) inside FP64 Newton / iterative refinement, Newton to 1e-8.
* ``mode="csr"``: FP64 residual, colored FP64 Hessian -> CSR, FP64 Jacobi-PCG.

Problems have the form of the large mixed-precision workloads, at sizes the oracle
finishes in seconds:

* Tet4: unit cube, 6 tets per cell (tests/test_sparse_tracer.py:29-70 split),
  interior nodes jittered by U(-0.2h, 0.2h) with seed 0, mu = 1, lambda = 10
  (tests/test_sparse_tracer.py:110-115), x = 0 clamped, u_x stepped on x = 1
  in increments of 0.03 (10 steps to a stretch of 1.3);
* Hex8: the same cube and loads, 2x2x2 Gauss, affine and jittered, increments
  of 0.06 (5 steps to 1.3).

Load steps follow the reference's runtime-value idiom: one Lifter with a
``RuntimeValue`` whose value is set per step through ``lifter.at[key].set``
(tatva/lifter/base.py:144,223-229).  Each step starts with the homogeneous
stretch predictor u_x += dx * x.  At every load step the oracle runs its own
Newton (colored FP64 Hessian, one JVP per color, Jacobi-PCG) from the GPU's
start state of that step with the same constraint values, and the test asserts

* the oracle's FP64 residual at the GPU state <= 1e-8 x the oracle residual at
  the step's start state (the metric's 1e-8 relative tolerance, evaluated by
  the oracle, not by the GPU's own residual kernel);
* the GPU state agrees with the oracle's converged state to 1e-6 relative
  (max |u_gpu - u_oracle| / max |u_oracle|): both stop at a relative residual,
  the states then differ by the conditioning times that.
"""

import numpy as np
import pytest

import tatva_amd as ta
from oracle import fem
from tatva_amd import energies, lifter, solvers, sparse

pytestmark = pytest.mark.gpu

NEWTON_TOL = 1e-8       # the metric's relative residual
STATE_RTOL = 1e-6       # GPU state vs the oracle's converged state


def to_np(t):
    return t.detach().cpu().numpy()


def stretch_case(kind, n, jitter):
    """Unit cube, x = 0 clamped, u_x prescribed on x = 1 through a RuntimeValue."""
    if kind == "tet4":
        mesh = ta.Mesh.unit_cube(n, type="tetrahedron", jitter=jitter, seed=0)
        el, oel = ta.Tetrahedron4(), fem.Tet4()
    else:
        mesh = ta.Mesh.unit_cube(n, jitter=jitter, seed=0)
        el, oel = ta.Hexahedron8(), fem.Hex8()
    op = ta.Operator(mesh, el)
    E = op.energy(energies.NeoHookean(1.0, 10.0))
    x0 = np.where(np.isclose(mesh.coords[:, 0], 0.0))[0]
    x1 = np.where(np.isclose(mesh.coords[:, 0], 1.0))[0]
    dofs = np.concatenate([(x0[:, None] * 3 + np.arange(3)).ravel(), x1 * 3])
    L = lifter.Lifter(mesh.n_nodes * 3, lifter.Fixed(dofs, lifter.RuntimeValue("load")))

    def values(d):
        return np.concatenate([np.zeros(x0.size * 3), np.full(x1.size, d)])

    prob = fem.Problem(oel, mesh.coords, mesh.elements, fem.NeoHookean(1.0, 10.0))
    return dict(mesh=mesh, op=op, E=E, L=L, values=values, dofs=dofs, prob=prob)


def oracle_pattern(C):
    mesh = C["mesh"]
    Lo = fem.Lifter(mesh.n_nodes * 3, fem.Fixed(C["dofs"]))
    So = Lo.adapt_sparsity(fem.pattern_from_mesh(mesh.elements, mesh.n_nodes, 3))
    return So, fem.distance2_colors(So.indptr, So.indices, So.shape[0])


def run_load_steps(C, mode, n_steps, dx):
    """GPU load-stepped Newton checked against the oracle at every step.
    Returns the per-step records (for the report of a failing assertion)."""
    import torch

    mesh, prob = C["mesh"], C["prob"]
    size = mesh.n_nodes * 3
    So, colors = oracle_pattern(C)
    kw = dict(lin_tol=1e-10) if mode == "csr" else dict(inner_tol=1e-3)
    solver = solvers.Newton(C["E"], C["L"].at["load"].set(C["values"](0.0)), mode=mode,
                            tol=NEWTON_TOL, chunk=16, **kw)
    ramp = torch.zeros(size, dtype=torch.float64, device="cuda")
    ramp[0::3] = torch.as_tensor(mesh.coords[:, 0], device="cuda")
    u = torch.zeros(size, dtype=torch.float64, device="cuda")
    records = []
    for s in range(1, n_steps + 1):
        u.add_(ramp, alpha=dx)                                   # predictor
        vals = C["values"](dx * s)
        Lo = fem.Lifter(size, fem.Fixed(C["dofs"], vals))
        start = to_np(u)
        start = Lo.lift(Lo.reduce(start), start)                 # same start, constraints set
        st = solver.solve(u, C["L"].at["load"].set(vals))
        assert st.converged, (mode, s, st.residual_history)
        ug = to_np(u)
        np.testing.assert_array_equal(ug[C["dofs"]], vals)
        uo, its, cg, hist = fem.newton(prob, Lo, start, So, colors, tol=1e-11, lin_tol=1e-12)
        r0 = np.linalg.norm(Lo.reduce_adjoint(prob.residual(start)))
        r_at_gpu = np.linalg.norm(Lo.reduce_adjoint(prob.residual(ug)))
        err = np.max(np.abs(ug - uo)) / np.max(np.abs(uo))
        rec = dict(step=s, newton=st.newton_iterations, cg=st.cg_iterations,
                   oracle_newton=its, oracle_rel_residual=r_at_gpu / r0, state_err=err)
        records.append(rec)
        assert hist[-1] <= 1e-10 * hist[0], ("oracle did not converge", rec)
        assert r_at_gpu <= NEWTON_TOL * r0, rec
        assert err <= STATE_RTOL, rec
    return records


CASES = [("tet4", 6, 0.2, 3, 0.03), ("hex8", 6, 0.0, 2, 0.06), ("hex8", 6, 0.2, 2, 0.06)]
IDS = ["tet4-jittered", "hex8-affine", "hex8-jittered"]


@pytest.mark.parametrize("mode", ["mixed", "csr"])
@pytest.mark.parametrize("kind,n,jitter,steps,dx", CASES, ids=IDS)
def test_load_stepped_neo_hookean_newton_matches_oracle(kind, n, jitter, steps, dx, mode):
    C = stretch_case(kind, n, jitter)
    recs = run_load_steps(C, mode, steps, dx)
    print(kind, jitter, mode, recs)
    if mode == "mixed":
        # the mixed path is Newton with an inexact (1e-3) inner solve: more outer steps
        assert all(r["newton"] >= 1 and r["cg"] > 0 for r in recs)


@pytest.mark.parametrize("kind,jitter", [("tet4", 0.2), ("hex8", 0.0), ("hex8", 0.2)],
                         ids=IDS)
def test_pcg_matfree_f32_matches_oracle_solve(kind, jitter):
    """tatva_pcg_matfree_f32 alone: the FP32 matrix-free Jacobi-PCG solution of
    H(u) x = b at a strained Neo-Hookean state, against the oracle's FP64
    Hessian (colored, one JVP per color) and a direct solve.  The state is the
    stretch predictor plus a 0.002 random perturbation: the Hessian is then
    positive definite (condition ~1e3; at 0.01 the jittered Tet4 one is not, and
    CG reports a breakdown).  Tolerance: the true FP64 relative residual <= 1e-4
    for an inner tolerance of 1e-6 (FP32 rounding of the operator times the
    conditioning), and x within 1e-2 of the direct solution (the conditioning
    times that); a wrong operator leaves O(1) errors.  The solve with the
    linearisation cached once and the one recomputing it in every product agree
    to FP32 rounding (iterations within 2, solutions within 1e-3)."""
    import scipy.sparse.linalg as spla
    import torch

    C = stretch_case(kind, 6, jitter)
    mesh, prob = C["mesh"], C["prob"]
    vals = C["values"](0.05)
    Lg = C["L"].at["load"].set(vals)
    Lo = fem.Lifter(mesh.n_nodes * 3, fem.Fixed(C["dofs"], vals))
    rng = np.random.default_rng(4)
    u = Lo.lift_from_zeros(0.002 * rng.standard_normal(Lo.size_reduced))
    u[0::3] += 0.05 * mesh.coords[:, 0]
    u = Lo.lift(Lo.reduce(u), u)
    So, colors = oracle_pattern(C)
    H = fem.hessian_csr(prob, Lo, u, So, colors)
    b = rng.standard_normal(Lo.size_reduced)
    x_ref = spla.spsolve(H.tocsc(), b)
    red = C["E"].reduced(Lg)
    u32 = torch.as_tensor(u, dtype=torch.float32, device="cuda")
    b32 = torch.as_tensor(b, dtype=torch.float32, device="cuda")
    solvers.set_linearization_cache("always")
    try:
        x32, info = solvers.pcg_matfree(red, u32, b32, tol=1e-6, chunk=16)
    finally:
        solvers.set_linearization_cache("never")
    assert info.converged, info
    x = x32.double().cpu().numpy()
    res = np.linalg.norm(H @ x - b) / np.linalg.norm(b)
    assert res <= 1e-4, res
    assert np.max(np.abs(x - x_ref)) <= 1e-2 * np.max(np.abs(x_ref))
    solvers.set_linearization_cache("never")
    try:
        x32b, info_b = solvers.pcg_matfree(red, u32, b32, tol=1e-6, chunk=16)
    finally:
        solvers.set_linearization_cache("never")
    # the same solve to FP32 rounding: the two kernels compute the same linearisation,
    # but the compiler fuses multiply-adds differently in the kernel that stores it
    # and in the one that uses it right away, so the last bits may differ
    xb = x32b.double().cpu().numpy()
    assert abs(info_b.iterations - info.iterations) <= 2, (info_b.iterations, info.iterations)
    assert np.linalg.norm(H @ xb - b) / np.linalg.norm(b) <= 1e-4
    assert np.max(np.abs(xb - x)) <= 1e-3 * np.max(np.abs(x))
    # the FP32 HVP is the FP32 restatement of the oracle's: same CG trajectory
    # up to rounding, so the iteration counts are close to the FP64 oracle PCG's
    _, k_or, _ = fem.pcg(H, b, tol=1e-6)
    assert abs(info.iterations - k_or) <= max(3, 0.1 * k_or), (info.iterations, k_or)


def test_colored_hessian_of_the_stretch_case_matches_oracle():
    """The FP64 path's colored Hessian on the jittered Tet4 stretch problem at a
    strained state: pattern and colors bit-exact, values to 1e-12."""
    C = stretch_case("tet4", 5, 0.2)
    mesh, prob = C["mesh"], C["prob"]
    vals = C["values"](0.03)
    Lg = C["L"].at["load"].set(vals)
    Lo = fem.Lifter(mesh.n_nodes * 3, fem.Fixed(C["dofs"], vals))
    u = Lo.lift_from_zeros(0.01 * np.random.default_rng(5).standard_normal(Lo.size_reduced))
    cm = sparse.ColoredMatrix.from_csr(Lg.adapt_sparsity(sparse.pattern_from_mesh(mesh, 3)))
    So, colors = oracle_pattern(C)
    np.testing.assert_array_equal(to_np(cm.indptr), So.indptr)
    np.testing.assert_array_equal(to_np(cm.indices), So.indices)
    np.testing.assert_array_equal(to_np(cm.colors), colors)
    K = sparse.jacfwd(C["E"].reduced(Lg).grad, cm)(Lo.reduce(u))
    Ko = fem.hessian_csr(prob, Lo, u, So, colors)
    rel = np.max(np.abs(to_np(K.data) - Ko.data)) / np.max(np.abs(Ko.data))
    assert rel <= 1e-12, rel


@pytest.mark.parametrize("lin", ["never", "always"])
def test_geometry_cache_gives_the_same_solve(lin):
    """With the affine Hex8 geometry cache (each element's J^-1 in FP32 and det J
    in FP64, written once per solve: tatva_set_option(TATVA_OPT_MATFREE_GEOM_CACHE,
    1)) the matrix-free FP32 PCG reads the values its products would compute from
    the coordinates: the solve is the same bit for bit (same iterations, same x),
    also with the linearisation written through the cache, and so is a mixed
    Newton load step of the stretch problem.  (The cache is on by default.)"""
    import torch

    C = stretch_case("hex8", 6, 0.0)
    E, L, mesh = C["E"], C["L"].at["load"].set(C["values"](0.05)), C["mesh"]
    nd = mesh.n_nodes * 3
    rng = np.random.default_rng(11)
    # a state with a positive definite Hessian: the stretch predictor and a small
    # perturbation
    u_np = 0.002 * rng.standard_normal(nd)
    u_np[0::3] += 0.05 * mesh.coords[:, 0]
    u = torch.as_tensor(u_np, dtype=torch.float32, device="cuda")
    red = E.reduced(L)
    b = torch.as_tensor(rng.standard_normal(red.size), dtype=torch.float32, device="cuda")
    runs, states = [], []
    solvers.set_linearization_cache(lin)
    try:
        for geo in (False, True):
            solvers.set_geometry_cache(geo)
            x, info = solvers.pcg_matfree(red, u, b, tol=1e-6, chunk=16)
            assert info.converged, (geo, info)
            runs.append((info.iterations, x.cpu().numpy()))
            uu = torch.zeros(nd, dtype=torch.float64, device="cuda")
            uu[0::3] = torch.as_tensor(0.05 * mesh.coords[:, 0], device="cuda")
            st = solvers.Newton(E, L, mode="mixed", tol=1e-8, inner_tol=1e-3, chunk=16).solve(uu)
            assert st.converged
            states.append((st.cg_iterations, uu.cpu().numpy()))
    finally:
        solvers.set_geometry_cache(True)   # the default
        solvers.set_linearization_cache("never")
    assert runs[0][0] == runs[1][0]
    np.testing.assert_array_equal(runs[1][1], runs[0][1])
    assert states[0][0] == states[1][0]
    np.testing.assert_array_equal(states[1][1], states[0][1])


@pytest.mark.parametrize("kind,jitter,profiling", [("hex8", 0.0, False), ("hex8", 0.1, True),
                                                   ("tet4", 0.2, False)])
def test_deferred_update_gives_the_same_solve(kind, jitter, profiling):
    """With each CG iteration's p and x updates applied by the next product's
    tile kernel (tatva_set_option(TATVA_OPT_MATFREE_DEFERRED_UPDATE, 1), the
    default) instead of by the iteration's step kernel, the matrix-free FP32
    PCG takes the same iterations to the same x bit for bit: converged, stopped
    by maxiter at an odd and an even count (the last update then reads either z
    buffer), with the sampled operator launch of the profiling mode, and through
    a mixed Newton load step.  n_free % 4 != 0 for the 6^3 cubes, so the
    one-entry tail of the 16-byte vector loops runs too."""
    import torch

    C = stretch_case(kind, 6, jitter)
    E, L, mesh = C["E"], C["L"].at["load"].set(C["values"](0.05)), C["mesh"]
    nd = mesh.n_nodes * 3
    rng = np.random.default_rng(13)
    u_np = 0.002 * rng.standard_normal(nd)
    u_np[0::3] += 0.05 * mesh.coords[:, 0]
    u = torch.as_tensor(u_np, dtype=torch.float32, device="cuda")
    red = E.reduced(L)
    if kind == "hex8":
        assert red.size % 4 != 0
    b = torch.as_tensor(rng.standard_normal(red.size), dtype=torch.float32, device="cuda")
    x0 = torch.as_tensor(0.1 * rng.standard_normal(red.size), dtype=torch.float32, device="cuda")
    runs = []
    solvers.set_profiling(profiling)
    try:
        for defer in (False, True):
            solvers.set_deferred_update(defer)
            out = []
            for maxiter, tol, x_start in ((None, 1e-6, None), (7, 1e-12, x0), (8, 1e-12, None)):
                x, info = solvers.pcg_matfree(red, u, b, x_start, tol=tol, maxiter=maxiter,
                                              chunk=4)
                assert info.converged == (maxiter is None), (defer, maxiter, info)
                out.append((info.iterations, x.cpu().numpy()))
            uu = torch.zeros(nd, dtype=torch.float64, device="cuda")
            uu[0::3] = torch.as_tensor(0.05 * mesh.coords[:, 0], device="cuda")
            st = solvers.Newton(E, L, mode="mixed", tol=1e-8, inner_tol=1e-3, chunk=16).solve(uu)
            assert st.converged
            out.append((st.cg_iterations, uu.cpu().numpy()))
            runs.append(out)
    finally:
        solvers.set_deferred_update(True)   # the default
        solvers.set_profiling(False)
    assert [r[0] for r in runs[0]][1:3] == [7, 8]
    for (it0, x0_), (it1, x1_) in zip(runs[0], runs[1]):
        assert it0 == it1
        np.testing.assert_array_equal(x1_, x0_)


@pytest.mark.parametrize("kind,jitter", [("tet4", 0.2), ("hex8", 0.0), ("hex8", 0.2)],
                         ids=IDS)
def test_full_vectors_solve_matches_oracle(kind, jitter):
    """The FP32 matrix-free PCG with Fixed constraints runs its CG on full-length
    vectors (tatva_set_option(TATVA_OPT_MATFREE_FULL_VECTORS, 1), the default:
    every dof of the mesh, the inverse diagonal 0 on the constrained ones, b, x
    and the diagonal scattered in and x gathered out through the dof map) or on
    the reduced vectors (0).  The products are the same; the dot products sum
    in another order.  Both solve H(u) x = b from a nonzero start to the true
    FP64 relative residual <= 1e-4 against the oracle's Hessian (inner tolerance
    1e-6), take iteration counts within 2 of each other and agree to 1e-3; a
    solve stopped by maxiter returns x of the same iteration either way (to FP32
    rounding of the dot order)."""
    import torch

    C = stretch_case(kind, 6, jitter)
    mesh, prob = C["mesh"], C["prob"]
    vals = C["values"](0.05)
    Lg = C["L"].at["load"].set(vals)
    Lo = fem.Lifter(mesh.n_nodes * 3, fem.Fixed(C["dofs"], vals))
    rng = np.random.default_rng(5)
    u = Lo.lift_from_zeros(0.002 * rng.standard_normal(Lo.size_reduced))
    u[0::3] += 0.05 * mesh.coords[:, 0]
    u = Lo.lift(Lo.reduce(u), u)
    So, colors = oracle_pattern(C)
    H = fem.hessian_csr(prob, Lo, u, So, colors)
    b = rng.standard_normal(Lo.size_reduced)
    x0 = 0.01 * rng.standard_normal(Lo.size_reduced)
    red = C["E"].reduced(Lg)
    u32 = torch.as_tensor(u, dtype=torch.float32, device="cuda")
    b32 = torch.as_tensor(b, dtype=torch.float32, device="cuda")
    x032 = torch.as_tensor(x0, dtype=torch.float32, device="cuda")
    out, short = [], []
    try:
        for full in (True, False):
            solvers.set_full_vectors(full)
            x32, info = solvers.pcg_matfree(red, u32, b32, x032, tol=1e-6, chunk=16)
            assert info.converged, (full, info)
            x = x32.double().cpu().numpy()
            res = np.linalg.norm(H @ x - b) / np.linalg.norm(b)
            assert res <= 1e-4, (full, res)
            out.append((info.iterations, x))
            xs, infos = solvers.pcg_matfree(red, u32, b32, x032, tol=1e-12, maxiter=5, chunk=4)
            assert infos.iterations == 5 and not infos.converged
            short.append(xs.double().cpu().numpy())
    finally:
        solvers.set_full_vectors(True)   # the default
    assert abs(out[0][0] - out[1][0]) <= 2, (out[0][0], out[1][0])
    assert np.max(np.abs(out[0][1] - out[1][1])) <= 1e-3 * np.max(np.abs(out[1][1]))
    assert np.max(np.abs(short[0] - short[1])) <= 1e-4 * np.max(np.abs(short[1]))
