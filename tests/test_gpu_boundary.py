"""GPU checks of the C-ABI's ownership and threading rules: caller-owned
workspaces (no library buffer outlives a call), reductions on two streams at
once, and caller-supplied colorings (tatva/sparse/base.py:59-74 with the
external colors of tests/test_sparse.py:65-67)."""

import ctypes

import numpy as np
import pytest

import tatva_amd as ta
from oracle import fem
from tatva_amd import _native as nat
from tatva_amd import energies, lifter, solvers, sparse

pytestmark = pytest.mark.gpu


def test_dot_on_two_streams_at_once():
    """tatva_dot takes its partial sums from the caller's workspace: two calls
    enqueued on two streams before either finishes return the values each
    gives alone (bit for bit) and the FP64 dot to 1e-14 relative."""
    import torch

    L = nat.lib()
    rng = np.random.default_rng(0)
    n = 3_000_001
    xs = [torch.as_tensor(rng.standard_normal(n), device="cuda") for _ in range(4)]
    alone = [float(solvers.dot(xs[0], xs[1])), float(solvers.dot(xs[2], xs[3]))]
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    outs = [torch.empty((), dtype=torch.float64, device="cuda") for _ in range(2)]
    wss = [torch.empty(int(L.tatva_dot_workspace(n)), dtype=torch.uint8, device="cuda")
           for _ in range(2)]
    torch.cuda.synchronize()
    for k in range(20):          # repeated, so that the two streams' kernels overlap
        for i, s in enumerate((s1, s2)):
            x, y = xs[2 * i], xs[2 * i + 1]
            nat.check(L.tatva_dot(n, nat.dptr(x), nat.dptr(y), nat.dptr(outs[i]),
                                  nat.dptr(wss[i]), wss[i].numel(), ctypes.c_void_p(s.cuda_stream)))
        torch.cuda.synchronize()
        got = [float(o) for o in outs]
        assert got == alone, (k, got, alone)
    for i in range(2):
        ref = float(np.dot(xs[2 * i].cpu().numpy(), xs[2 * i + 1].cpu().numpy()))
        assert abs(alone[i] - ref) <= 1e-14 * np.sqrt(n) * max(abs(ref), 1.0)
    with pytest.raises(ValueError, match="workspace too small"):
        nat.check(L.tatva_dot(n, nat.dptr(xs[0]), nat.dptr(xs[1]), nat.dptr(outs[0]),
                              nat.dptr(wss[0]), 8, nat.stream()))


def test_pcg_csr_keeps_no_device_memory():
    """tatva_pcg_csr without a block view lays the matrix out as sliced ELL in
    the caller's workspace: after the solve and the workspace are released,
    the device memory in use is back at its level before the solve (a library
    buffer of the copy would hold 12 bytes per non-zero)."""
    import torch

    m = ta.Mesh.unit_square(96, 96)
    op = ta.Operator(m, ta.Tri3())
    E = op.energy(energies.Poisson(1.0))
    bnd = np.where(np.isclose(m.coords, 0).any(1) | np.isclose(m.coords, 1).any(1))[0]
    L = lifter.Lifter(m.n_nodes, lifter.Fixed(bnd, 0.0))
    cm = sparse.ColoredMatrix.from_csr(L.adapt_sparsity(sparse.pattern_from_mesh(m, 1)))
    K = sparse.jacfwd(E.reduced(L).grad, cm)(np.zeros(L.size_reduced))
    b = torch.ones(L.size_reduced, dtype=torch.float64, device="cuda")
    import gc

    solvers.pcg(K, b, tol=1e-10, workspace=solvers._Workspace())  # graphs, kernels loaded

    def free_now():
        gc.collect()      # tensors of earlier tests still waiting for the collector
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        return torch.cuda.mem_get_info()[0]

    free0 = free_now()
    ws = solvers._Workspace()
    x, info = solvers.pcg(K, b, tol=1e-10, workspace=ws)
    assert info.converged
    assert ws.buf.numel() >= 12 * cm.nnz   # the sliced ELL copy is in it
    del ws, x
    free1 = free_now()
    assert free1 >= free0 - 2 * 2 ** 20, (free0 - free1)
    # a workspace sized for the block view only is too small for the ELL copy
    Lb = nat.lib()
    n = K.shape[0]
    small = torch.empty(int(Lb.tatva_pcg_csr_workspace(n, None)), dtype=torch.uint8,
                        device="cuda")
    xs = torch.zeros(n, dtype=torch.float64, device="cuda")
    st = nat.PcgStats()
    with pytest.raises(ValueError, match="too small for the sliced ELL copy"):
        nat.check(Lb.tatva_pcg_csr(n, K.nnz, nat.dptr(K.indptr), nat.dptr(K.indices),
                                   nat.dptr(K.data), None, 1, nat.dptr(b), nat.dptr(xs), 1e-10,
                                   100, 16, nat.dptr(small), small.numel(), ctypes.byref(st),
                                   nat.stream()))


def test_caller_coloring_ldf_gives_the_greedy_data():
    """ColoredMatrix.from_csr with a caller's valid, non-greedy coloring (the
    reference's largest-degree-first one, restated in the oracle) decompresses
    to the same Hessian as the greedy coloring (1e-12) and the dense oracle
    Hessian; an invalid coloring raises ValueError."""
    m = ta.Mesh.unit_square(7, 6)
    op = ta.Operator(m, ta.Tri3())
    E = op.energy(energies.LinearElasticity(1.0, 0.0))
    S = sparse.pattern_from_mesh(m, 2)
    n = S.shape[0]
    ldf = fem.largest_degree_first_distance2_colors(S.indptr, S.indices, n)
    cm_g = sparse.ColoredMatrix.from_csr(S)
    cm_l = sparse.ColoredMatrix.from_csr(S, colors=ldf)
    assert not np.array_equal(cm_g.colors.cpu().numpy(), ldf)
    u = 0.01 * np.random.default_rng(1).standard_normal(n)
    Kg = sparse.jacfwd(E.grad, cm_g)(u)
    Kl = sparse.jacfwd(E.grad, cm_l)(u)
    dg, dl = Kg.data.cpu().numpy(), Kl.data.cpu().numpy()
    assert np.max(np.abs(dl - dg)) <= 1e-12 * np.max(np.abs(dg))
    prob = fem.Problem(fem.Tri3(), m.coords, m.elements, fem.LinearElasticity(1.0, 0.0))
    H = np.stack([prob.hvp(u, e) for e in np.eye(n)], axis=1)
    assert np.max(np.abs(Kl.to_dense() - H)) <= 1e-12 * np.max(np.abs(H))
    bad = ldf.copy()
    bad[S.indices[S.indptr[5] + 1]] = bad[S.indices[S.indptr[5]]]
    with pytest.raises(ValueError, match="not a distance-2 coloring"):
        sparse.ColoredMatrix.from_csr(S, colors=bad)
