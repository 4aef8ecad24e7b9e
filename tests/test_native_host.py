"""CPU tests of the native library: it loads, exports every declared symbol, and
its host-side setup (pattern, coloring, constrained submatrix) matches the
oracle bit for bit.  Also the host-side API surface (validation, errors)."""

import os
import re

import numpy as np
import pytest
import scipy.sparse as sps

import tatva_amd as ta
from oracle import fem
from tatva_amd import _native as nat
from tatva_amd import lifter, sparse

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tatva_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tatva_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = nat.lib()
    syms = declared_symbols()
    assert len(syms) >= 35
    for s in syms:
        assert hasattr(L, s), s
        assert s in nat.EXPORTED, f"{s} has no ctypes signature"
    assert L.tatva_version() >= 1000


def test_library_binds_to_the_hip_runtime_torch_maps():
    """Loaded first thing in a fresh process, the library must not map a
    second HIP/HSA runtime beside torch's (that second runtime sees no GPU)."""
    import subprocess
    import sys
    code = ("import os\n"
            "from tatva_amd import _native as nat\n"
            "nat.lib()\n"
            "maps = open(f'/proc/{os.getpid()}/maps').read().splitlines()\n"
            "for key in ('amdhip64', 'hsa-runtime64'):\n"
            "    print(key, len({l.split()[-1] for l in maps if key in l}))\n")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    assert "amdhip64 1" in out.stdout and "hsa-runtime64 1" in out.stdout, out.stdout


def test_error_plumbing():
    h = nat.c_void()
    rc = nat.lib().tatva_pattern_from_mesh(None, 1, 3, 3, 1, None, 3, nat.ctypes.byref(h))
    assert rc == nat.EINVAL
    with pytest.raises(ValueError, match="invalid arguments"):
        nat.check(rc)


MESHES = {
    "tri3_9x7": (lambda: ta.Mesh.unit_square(9, 7), lambda: fem.unit_square_tri3(9, 7), 2),
    "tri3_poisson": (lambda: ta.Mesh.unit_square(12, 12), lambda: fem.unit_square_tri3(12, 12), 1),
    "hex8_7": (lambda: ta.Mesh.unit_cube(7), lambda: fem.cube_hex8(7), 3),
    "tet4_jit": (lambda: ta.Mesh.box((4, 3, 5), type="tetrahedron", jitter=0.2, seed=0),
                 lambda: fem.box_tet4(4, 3, 5, jitter=0.2, seed=0), 3),
}


@pytest.mark.parametrize("name", list(MESHES))
def test_generators_match_oracle(name):
    mk, mko, _ = MESHES[name]
    m = mk()
    X, E = mko()
    np.testing.assert_array_equal(m.elements, E)
    np.testing.assert_array_equal(m.coords, X)


@pytest.mark.parametrize("name", list(MESHES))
def test_pattern_coloring_reduction_match_oracle(name):
    mk, _, ndpn = MESHES[name]
    m = mk()
    S = sparse.pattern_from_mesh(m, ndpn)
    So = fem.pattern_from_mesh(m.elements, m.n_nodes, ndpn)
    np.testing.assert_array_equal(S.indptr, So.indptr)
    np.testing.assert_array_equal(S.indices, So.indices)
    assert S.dtype == np.int8 and np.all(S.data == 1)
    c = sparse.distance2_colors(S.indptr, S.indices, S.shape[0])
    np.testing.assert_array_equal(c, fem.distance2_colors(So.indptr, So.indices, So.shape[0]))
    # constrain the x=0 face fully and one component on the x=1 face (breaks node runs)
    x0 = np.where(np.isclose(m.coords[:, 0], 0.0))[0]
    x1 = np.where(np.isclose(m.coords[:, 0], 1.0))[0]
    dofs = np.concatenate([(x0[:, None] * ndpn + np.arange(ndpn)).ravel(), x1 * ndpn])
    L = lifter.Lifter(S.shape[0], lifter.Fixed(dofs))
    Lo = fem.Lifter(S.shape[0], fem.Fixed(dofs))
    np.testing.assert_array_equal(L.free_dofs, Lo.free)
    R = L.adapt_sparsity(S)
    Ro = Lo.adapt_sparsity(So)
    Ro.sort_indices()
    np.testing.assert_array_equal(R.indptr, Ro.indptr)
    np.testing.assert_array_equal(R.indices, Ro.indices)
    cr = sparse.distance2_colors(R.indptr, R.indices, R.shape[0])
    np.testing.assert_array_equal(cr, fem.distance2_colors(Ro.indptr, Ro.indices, Ro.shape[0]))


def test_coloring_general_patterns():
    """Patterns without FEM block structure: missing diagonals, asymmetric rows."""
    rng = np.random.default_rng(3)
    for n, dens in [(60, 0.05), (200, 0.02), (500, 0.004)]:
        A = sps.random(n, n, density=dens, random_state=rng, format="csr")
        A.data[:] = 1
        A.sort_indices()
        ip = A.indptr.astype(np.int32)
        ix = A.indices.astype(np.int32)
        np.testing.assert_array_equal(sparse.distance2_colors(ip, ix, n),
                                      fem.distance2_colors_py(ip, ix, n))
    # identical consecutive rows with diagonal present: exercises the run grouping
    blocks = sps.kron(sps.random(40, 40, density=0.08, random_state=rng) + sps.eye(40),
                      np.ones((3, 3)), format="csr")
    blocks.data[:] = 1
    blocks.sort_indices()
    ip, ix = blocks.indptr.astype(np.int32), blocks.indices.astype(np.int32)
    np.testing.assert_array_equal(sparse.distance2_colors(ip, ix, 120),
                                  fem.distance2_colors_py(ip, ix, 120))


def test_pattern_with_dof_map_equals_submatrix():
    m = ta.Mesh.unit_cube(5)
    n = m.n_nodes * 3
    keep = np.ones(n, dtype=bool)
    keep[np.random.default_rng(0).choice(n, 40, replace=False)] = False
    dof_map = np.full(n, -1, dtype=np.int32)
    dof_map[keep] = np.arange(keep.sum(), dtype=np.int32)
    h = nat.c_void()
    L = nat.lib()
    el = np.ascontiguousarray(m.elements)
    nat.check(L.tatva_pattern_from_mesh(el.ctypes.data, el.shape[0], 8, m.n_nodes, 3,
                                        dof_map.ctypes.data, int(keep.sum()), nat.ctypes.byref(h)))
    nnz = L.tatva_csr_host_nnz(h)
    ip = np.empty(int(keep.sum()) + 1, dtype=np.int32)
    ix = np.empty(nnz, dtype=np.int32)
    nat.check(L.tatva_csr_host_copy(h, ip.ctypes.data, ix.ctypes.data))
    L.tatva_csr_host_free(h)
    So = fem.pattern_from_mesh(m.elements, m.n_nodes, 3)[keep][:, keep]
    So.sort_indices()
    np.testing.assert_array_equal(ip, So.indptr)
    np.testing.assert_array_equal(ix, So.indices)


def test_operator_validation_errors():
    X = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    with pytest.raises(ValueError, match="outside"):
        ta.Operator(ta.Mesh(X, np.array([[0, 1, 3]])), ta.Tri3())
    with pytest.raises(ValueError, match="negative"):
        ta.Operator(ta.Mesh(X, np.array([[0, -1, 2]])), ta.Tri3())
    with pytest.raises(TypeError):
        ta.Mesh(X, np.array([[0.0, 1.0, 2.0]]))
    with pytest.raises(ValueError, match="at least one element"):
        ta.Operator(ta.Mesh(X, np.zeros((0, 3), dtype=np.int32)), ta.Tri3())
    with pytest.raises(ValueError, match="3-D"):
        ta.Operator(ta.Mesh(X, np.array([[0, 1, 2, 0]])), ta.Tetrahedron4())
    op = ta.Operator(ta.Mesh(X, np.array([[0, 1, 2]])), ta.Tri3())
    assert op.batch_size == 1
    with pytest.raises(NotImplementedError):
        op.map(lambda xi, v: v)
    with pytest.raises(NotImplementedError):
        op.map_over_elements(lambda v: v)


def test_elements_and_quadrature():
    assert ta.Tri3() == ta.Tri3() and hash(ta.Hexahedron8()) == hash(ta.Hexahedron8())
    h = ta.Hexahedron8()
    np.testing.assert_allclose(np.abs(h.quad_points), 1 / np.sqrt(3))
    assert h.quad_weights.sum() == 8.0
    t = ta.Tetrahedron4(quad_points=[[0.1, 0.2, 0.3]], quad_weights=[1 / 6])
    assert t != ta.Tetrahedron4()
    with pytest.raises(ValueError):
        ta.Tri3(quad_points=np.zeros((28, 2)), quad_weights=np.ones(28))


def test_lifter_host_logic():
    L = lifter.Lifter(6, lifter.Fixed(np.array([0, 5])), lifter.Fixed([2], lifter.RuntimeValue("t")))
    np.testing.assert_array_equal(L.free_dofs, [1, 3, 4])
    np.testing.assert_array_equal(L.constrained_dofs, [0, 2, 5])
    assert L.size_reduced == 3
    np.testing.assert_array_equal(L.dof_map, [-1, 0, -1, 1, 2, -1])
    L2 = L.at["t"].set(3.0)
    assert L2._runtime_values == {"t": 3.0} and L._runtime_values == {}
    assert L2 != L and L2 == L.with_values({"t": 3.0})
    with pytest.raises(lifter.LifterError):
        L.with_values({"nope": 1.0})
    with pytest.raises(lifter.LifterError):
        L.constraints[1].resolved(L._runtime_values)
    assert isinstance(hash(L), int)
    with pytest.raises(lifter.LifterError):
        lifter.lifted(lambda u: u)(L, np.zeros(2))
    with pytest.raises(ImportError):
        sparse.create_sparsity_pattern  # noqa: B018


def test_jacfwd_rejects_callables():
    X = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    S = sparse.pattern_from_mesh(ta.Mesh(X, np.array([[0, 1, 2]])), 1)
    cm_cls = sparse.ColoredMatrix
    with pytest.raises(NotImplementedError):
        sparse._HessianAssembler(lambda u: u, cm_cls(None, None, None, S.shape, None, 3))
    with pytest.raises(NotImplementedError):
        ta.grad(lambda u: u)


def test_periodic_layout_and_sparsity_match_oracle():
    """Read map / reduce rows of Fixed + Periodic, and augment + reduce of the
    reference's periodic sparsity test layout (tests/test_sparse_tracer.py:162-216)."""
    L = lifter.Lifter(6, lifter.Fixed([0, 5]), lifter.Periodic([2], [1]))
    np.testing.assert_array_equal(L.free_dofs, [1, 3, 4])
    np.testing.assert_array_equal(L.read_map, [-1, 0, 0, 1, 2, -1])
    np.testing.assert_array_equal(L._layout.red_ptr, [0, 2, 3, 4])
    np.testing.assert_array_equal(L._layout.red_list, [1, 2, 3, 4])
    # order matters: a periodic slave copying a master fixed later reads nothing
    L2 = lifter.Lifter(4, lifter.Periodic([2], [1]), lifter.Fixed([1]))
    np.testing.assert_array_equal(L2.read_map, [0, -1, -1, 1])
    m = ta.Mesh.unit_square(4, 4)
    X = m.coords
    bot = np.where(X[:, 1] == 0.0)[0]
    left = np.where(X[:, 0] == 0.0)[0]
    left = left[np.argsort(X[left, 1])]
    right = np.where(X[:, 0] == 1.0)[0]
    right = right[np.argsort(X[right, 1])]
    c = np.arange(2)
    cons = [((bot[:, None] * 2 + c).ravel(),),
            ((left[:, None] * 2 + c).ravel(), (right[:, None] * 2 + c).ravel())]
    L = lifter.Lifter(m.n_nodes * 2, lifter.Fixed(*cons[0]), lifter.Periodic(*cons[1]))
    Lo = fem.Lifter(m.n_nodes * 2, fem.Fixed(*cons[0]), fem.Periodic(*cons[1]))
    S = sparse.pattern_from_mesh(m, 2)
    So = fem.pattern_from_mesh(m.elements, m.n_nodes, 2)
    aug, aug_o = L.augment_sparsity(S), fem.Periodic(*cons[1]).augment(So)
    assert (aug != aug_o).nnz == 0
    R, Ro = L.adapt_sparsity(S), Lo.adapt_sparsity(So)
    np.testing.assert_array_equal(R.indptr, Ro.indptr)
    np.testing.assert_array_equal(R.indices, Ro.indices)
    # read map = linear part of the oracle lift, column by column
    for j in range(0, L.size_reduced, 7):
        e = np.zeros(L.size_reduced)
        e[j] = 1.0
        np.testing.assert_array_equal(np.nonzero(Lo.lift_linear(e))[0],
                                      np.nonzero(L.read_map == j)[0])


def test_field_patterns_match_oracle():
    """Displacement (2 dofs, all nodes), a scalar on a node subset and two
    global dofs (pattern_from_compound's nodal + diagonal cases)."""
    m = ta.Mesh.unit_square(5, 4)
    N = m.n_nodes
    u = np.arange(2 * N).reshape(N, 2)
    sub = np.where(m.coords[:, 0] <= 0.4)[0]
    p = np.full((N, 1), -1)
    p[sub, 0] = 2 * N + np.arange(sub.size)
    glob = np.array([2 * N + sub.size, 2 * N + sub.size + 1])
    size = int(glob[-1]) + 1
    S = sparse.pattern_from_fields(m, size, nodal=[u, p], diagonal=[glob])
    ed = np.concatenate([u[m.elements].reshape(m.n_elements, -1),
                         p[m.elements].reshape(m.n_elements, -1)], axis=1)
    So = fem.pattern_from_element_dofs(ed, size, glob)
    np.testing.assert_array_equal(S.indptr, So.indptr)
    np.testing.assert_array_equal(S.indices, So.indices)
    blocks = sparse.pattern_from_fields(m, size, nodal=[u, p], diagonal=[glob],
                                        block_slices=[slice(0, 2 * N), slice(2 * N, size)])
    assert blocks[0][1].shape == (2 * N, size - 2 * N)
    assert (blocks[1][0] != So[2 * N:, :2 * N]).nnz == 0
    # empty element table: diagonal only
    D = sparse.pattern_from_element_dofs(np.zeros((0, 3), dtype=np.int32), 4, [1, 2])
    np.testing.assert_array_equal(D.indptr, [0, 0, 1, 2, 2])
    with pytest.raises(ValueError):
        sparse.pattern_from_element_dofs(np.array([[0, 7]]), 5)


def test_jittered_box_keeps_every_element_positive():
    """C3's jittered Tet4 box must not invert an element (every signed volume,
    and every Hex8 corner Jacobian, positive); a jitter that inverts raises."""
    import tatva_amd as ta
    from tatva_amd.mesh import ElementType, _min_corner_volume

    for kind, et in (("tetrahedron", ElementType.TETRAHEDRON), ("hexahedron", ElementType.HEXAHEDRON)):
        m = ta.Mesh.box(6, type=kind, jitter=0.2, seed=0)
        assert _min_corner_volume(m.coords, m.elements, et) > 0.0
        plain = ta.Mesh.box(6, type=kind)
        assert np.isclose(_min_corner_volume(plain.coords, plain.elements, et), (1 / 6) ** 3)
    with pytest.raises(ValueError):
        ta.Mesh.box(8, type="tetrahedron", jitter=0.9, seed=0)


def test_newton_rejects_a_lifter_with_another_layout():
    """Newton.solve keeps the reduced sizes and pattern of its constructor's
    lifter: new constraint values pass, other constrained dofs raise before any
    kernel could read or write past the end of a reduced buffer."""
    import types

    from tatva_amd import solvers

    L = lifter.Lifter(12, lifter.Fixed([0, 1, 2, 9], lifter.RuntimeValue("s")))
    nw = solvers.Newton.__new__(solvers.Newton)
    nw.red = types.SimpleNamespace(lifter=L.at["s"].set(np.zeros(4)))
    nw._check_layout(L.at["s"].set(np.ones(4)))                     # same layout object
    nw._check_layout(lifter.Lifter(12, lifter.Fixed([0, 1, 2, 9], 0.5)))  # equal index sets
    for other in (lifter.Lifter(12, lifter.Fixed([0, 1, 2], 0.0)),
                  lifter.Lifter(12, lifter.Fixed([0, 1, 2, 10], 0.0)),
                  lifter.Lifter(12, lifter.Fixed([0, 1, 2], 0.0), lifter.Periodic([9], [3])),
                  lifter.Lifter(15, lifter.Fixed([0, 1, 2, 9], 0.0))):
        with pytest.raises(ValueError, match="constrains other dofs"):
            nw._check_layout(other)


def test_distance2_color_check_matches_the_oracle():
    """tatva_check_distance2_colors (the validation of caller colors in
    ColoredMatrix.from_csr) against the oracle's check: the greedy and the
    largest-degree-first colorings pass, a coloring with one color changed to
    that of another column of the same row is reported at the first such row,
    and from_csr raises ValueError for it before any upload."""
    import ctypes

    X, el = fem.unit_square_tri3(6, 5)
    S = fem.pattern_from_mesh(el, X.shape[0], 2)
    n = S.shape[0]
    ip = np.ascontiguousarray(S.indptr, dtype=np.int32)
    ix = np.ascontiguousarray(S.indices, dtype=np.int32)

    def check(colors):
        c = np.ascontiguousarray(colors, dtype=np.int32)
        bad = ctypes.c_int64(0)
        nat.check(nat.lib().tatva_check_distance2_colors(n, ip.ctypes.data, ix.ctypes.data,
                                                          c.ctypes.data, ctypes.byref(bad)))
        return bad.value

    greedy = fem.distance2_colors(S.indptr, S.indices, n)
    ldf = fem.largest_degree_first_distance2_colors(S.indptr, S.indices, n)
    assert check(greedy) == -1 and check(ldf) == -1
    bad = greedy.copy()
    r = 7
    bad[S.indices[S.indptr[r] + 2]] = bad[S.indices[S.indptr[r]]]
    first = next(i for i in range(n) if not fem.is_distance2_coloring(
        S.indptr[i:i + 2] - S.indptr[i], S.indices[S.indptr[i]:S.indptr[i + 1]], bad))
    assert check(bad) == first
    with pytest.raises(ValueError, match="not a distance-2 coloring"):
        sparse.ColoredMatrix.from_csr(S, colors=bad)


def test_padded_node_layout_matches_brute_force():
    """The whole-node layout of the FP64 Newton system (u_x prescribed on x = 1,
    x = 0 clamped): its pattern is node-blocked and is the pattern of the dofs of
    every node with a free dof; the zeroed positions are exactly the
    off-diagonal entries in rows or columns of constrained dofs and the unit
    ones their diagonal (brute force over the pattern's entries); each reduced
    dof maps to its own dof."""
    from tatva_amd.solvers import padded_node_layout

    m = ta.Mesh.unit_cube(4, type="tetrahedron", jitter=0.2, seed=0)
    x0 = np.where(np.isclose(m.coords[:, 0], 0))[0]
    x1 = np.where(np.isclose(m.coords[:, 0], 1))[0]
    dofs = np.concatenate([(x0[:, None] * 3 + np.arange(3)).ravel(), x1 * 3])
    L = lifter.Lifter(m.n_nodes * 3, lifter.Fixed(dofs, 0.0))
    L_pad, S, red_to_pad, zero_pos, diag_pos = padded_node_layout(m, L, 3)
    pad_dofs = L_pad.free_dofs
    np.testing.assert_array_equal(np.unique(pad_dofs // 3), np.unique(L.free_dofs // 3))
    assert pad_dofs.size == 3 * np.unique(L.free_dofs // 3).size
    np.testing.assert_array_equal(pad_dofs[red_to_pad], L.free_dofs)
    ip, ix = S.indptr, S.indices
    for k in range(S.shape[0] // 3):
        for j in (1, 2):
            np.testing.assert_array_equal(ix[ip[3 * k]:ip[3 * k + 1]],
                                          ix[ip[3 * k + j]:ip[3 * k + j + 1]])
    So = fem.Lifter(L.size, fem.Fixed(np.setdiff1d(np.arange(L.size), pad_dofs))).adapt_sparsity(
        fem.pattern_from_mesh(m.elements, m.n_nodes, 3))
    np.testing.assert_array_equal(ip, So.indptr)
    np.testing.assert_array_equal(ix, So.indices)
    con = set(np.nonzero(~np.isin(pad_dofs, L.free_dofs))[0].tolist())
    rows = np.repeat(np.arange(S.shape[0]), np.diff(ip))
    want_zero = [k for k, (r, c) in enumerate(zip(rows, ix)) if r != c and (r in con or c in con)]
    want_diag = [k for k, (r, c) in enumerate(zip(rows, ix)) if r == c and r in con]
    np.testing.assert_array_equal(zero_pos, want_zero)
    np.testing.assert_array_equal(np.sort(diag_pos), want_diag)
    # nothing to pad when whole nodes are constrained
    assert padded_node_layout(m, lifter.Lifter(m.n_nodes * 3, lifter.Fixed(
        (x0[:, None] * 3 + np.arange(3)).ravel(), 0.0)), 3) is None


def test_pcg_workspace_must_be_aligned():
    """The PCG entry points move their vectors in 16-byte loads from 256-byte
    offsets of the workspace, so they reject a workspace that is not 256-byte
    aligned before any device work (the pointers below are never dereferenced)."""
    L = nat.lib()
    n = 1000
    need = int(L.tatva_pcg_csr_workspace(n, None))
    fake = nat.ctypes.c_void_p(1 << 20)
    st = nat.PcgStats()
    rc = L.tatva_pcg_csr(n, 5 * n, fake, fake, fake, None, 1, fake, fake, 1e-8, 10, 16,
                         nat.ctypes.c_void_p((1 << 20) + 16), need, nat.ctypes.byref(st), None)
    assert rc == nat.EINVAL
    with pytest.raises(ValueError, match="256-byte aligned"):
        nat.check(rc)


def test_solver_options_take_their_documented_values_only():
    """tatva_set_option (include/tatva_amd.h) accepts each option's documented
    values and rejects others, and an unknown option, with TATVA_EINVAL before
    anything runs; the Python setters mirror it and leave the defaults as they
    found them."""
    from tatva_amd import solvers

    L = nat.lib()
    valid = {solvers.OPT_MATFREE_LIN_CACHE: (0, 1, 2), solvers.OPT_MATFREE_GEOM_CACHE: (0, 1),
             solvers.OPT_MATFREE_DEFERRED_UPDATE: (0, 1), solvers.OPT_MATFREE_FULL_VECTORS: (0, 1)}
    defaults = {solvers.OPT_MATFREE_LIN_CACHE: 0, solvers.OPT_MATFREE_GEOM_CACHE: 1,
                solvers.OPT_MATFREE_DEFERRED_UPDATE: 1, solvers.OPT_MATFREE_FULL_VECTORS: 1}
    try:
        for opt, values in valid.items():
            for v in values:
                assert L.tatva_set_option(opt, v) == nat.OK, (opt, v)
            for v in (-1, max(values) + 1):
                assert L.tatva_set_option(opt, v) == nat.EINVAL, (opt, v)
        assert L.tatva_set_option(99, 0) == nat.EINVAL
        with pytest.raises(ValueError, match="unknown option"):
            nat.check(L.tatva_set_option(99, 0))
        with pytest.raises(ValueError):
            solvers.set_linearization_cache("sometimes")
        solvers.set_deferred_update(False)
        solvers.set_full_vectors(False)
    finally:
        for opt, v in defaults.items():
            assert L.tatva_set_option(opt, v) == nat.OK


def test_hvp_tile_bytes_counts_every_array_once():
    """bench.hvp_tile_bytes (the roofline's algorithmic bytes) against a count
    written per array: Hex8 and Tet4 tiles read incidence lists (a uint16 entry
    per element node, a uint16 end per slot) and one color offset per tile;
    Tri3 tiles read every color offset."""
    import bench

    class Op:
        def __init__(self, kind, lay):
            self.element = type("El", (), {"kind": kind})()
            self._lay = lay

        def device_layout(self):
            return dict(self._lay)

    E, S, T, C, n = 4096, 7300, 16, 12, 12_000
    lay = {"n_elements": E, "n_slots": S, "n_tiles": T, "max_colors": C, "n_quad": 8}
    # Hex8, geometry cache, full-length vectors, deferred update
    hex8 = (2 * 8 * E          # local node numbers
            + 44 * E           # geometry cache entry
            + S * (4 + 12 + 12 + 12 + 4)  # node id, state, tangent, partial sums, position
            + 2 * 8 * E + 2 * S  # incidence entries and list ends
            + T * (2 + 8)      # element count, slot range
            + 20 * n)          # deferred p, x update
    assert bench.hvp_tile_bytes(Op(2, lay), geo=True, defer_n=n, mapped=False)[0] == hex8
    # Tet4, coordinates read (no cache), reduced vectors through the dof map
    tet4 = (2 * 4 * E + S * (4 + 8 * 3 + 12 + 12 + 12 + 12 + 4) + 2 * 4 * E + 2 * S
            + T * (2 + 8))
    assert bench.hvp_tile_bytes(Op(1, lay), geo=False, mapped=True)[0] == tet4
    # Tri3, two components: every color offset, 2-D coordinates
    tri3 = (2 * 3 * E + S * (4 + 8 * 2 + 8 + 8 + 8 + 8 + 4) + T * (2 * (C + 1) + 8))
    assert bench.hvp_tile_bytes(Op(0, lay), dens_ncomp=2, mapped=True)[0] == tri3
