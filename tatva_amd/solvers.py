"""Jacobi-PCG and Newton on the GPU.

tatva has no solver of its own (its only linear solve is ``spsolve`` inside the
L2 projection, tatva/utils.py:186-198); users write Newton/CG around
``sparse.jacfwd`` or ``jax.jvp``.  This module provides that loop natively:

* ``pcg(A, b)`` -- FP64 Jacobi-PCG on a ``ColoredMatrix`` (CSR SpMV fused with
  p.Ap), iterations run as hipGraph-captured chunks with a device-side
  convergence flag, so the host synchronises once per chunk.
* ``pcg_matfree(E_red, u_full, b)`` -- FP32 Jacobi-PCG whose operator is the
  matrix-free FP32 Hessian-vector product of a built-in energy.
* ``Newton`` -- Newton on the free dofs of a ``Lifter``.  ``mode="csr"``:
  FP64 residual, colored FP64 Hessian -> CSR, FP64 PCG.  ``mode="mixed"``:
  FP64 residual and energy, FP32 matrix-free PCG for the correction, the outer
  Newton/iterative-refinement update in FP64.
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field, replace

import numpy as np

from . import _native as nat
from .energies import Energy, ReducedEnergy
from .sparse import ColoredMatrix, _HessianAssembler


@dataclass
class PcgInfo:
    iterations: int
    rel_residual: float
    status: int
    b_norm: float
    op_ms: float = 0.0        # profiling: summed duration of sampled operator launches
    op_samples: int = 0
    op_main_ms: float = 0.0   # profiling: the same launches' main kernel (SpMV / HVP tile kernel)

    @property
    def converged(self) -> bool:
        return self.status == nat.OK


def _stats_to_info(st: nat.PcgStats) -> PcgInfo:
    return PcgInfo(int(st.iterations), float(st.rel_residual), int(st.status), float(st.b_norm),
                   float(st.op_ms), int(st.op_samples), float(st.op_main_ms))


class _Workspace:
    def __init__(self):
        self.buf = None

    def get(self, nbytes: int):
        t = nat.torch()
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = t.empty(int(nbytes), dtype=t.uint8, device=nat.device())
        return self.buf


_WS_CSR = _Workspace()


def pcg(A: ColoredMatrix, b, x0=None, *, tol: float = 1e-10, maxiter: int | None = None,
        chunk: int = 32, workspace=None, raise_on_failure: bool = False,
        block_size: int | None = None):
    """Solve A x = b (FP64, Jacobi preconditioner); returns (x, PcgInfo).

    Stops when ||r_k|| <= tol ||b||.  With ``block_size`` (dofs per node) the
    operator reads the matrix through its node-block view when the pattern
    allows it (same arithmetic per row, fewer index bytes).
    """
    t = nat.torch()
    n = A.shape[0]
    bv = nat.to_device(b).reshape(-1).contiguous()
    x = t.zeros(n, dtype=t.float64, device=bv.device) if x0 is None else \
        nat.to_device(x0).reshape(-1).clone()
    maxiter = 10 * n if maxiter is None else int(maxiter)
    L = nat.lib()
    view = A.block_view(block_size) if block_size else None
    bc, bs = (view[0], view[1]) if view is not None else (None, 1)
    key = ("pcg_csr_workspace", view is None)
    nbytes = A._cache.get(key)
    if nbytes is None:
        # without a block view the workspace also holds the sliced-ELL copy
        hip = None if view is not None else np.ascontiguousarray(A.host_indptr(), dtype=np.int32)
        nbytes = int(L.tatva_pcg_csr_workspace(n, None if hip is None else hip.ctypes.data))
        A._cache[key] = nbytes
    ws = (workspace or _WS_CSR).get(nbytes)
    st = nat.PcgStats()
    nat.check(L.tatva_pcg_csr(n, A.nnz, nat.dptr(A.indptr), nat.dptr(A.indices), nat.dptr(A.data),
                              nat.dptr(bc), bs, nat.dptr(bv), nat.dptr(x), float(tol), maxiter,
                              int(chunk), nat.dptr(ws), ws.numel(), ctypes.byref(st),
                              nat.stream()))
    info = _stats_to_info(st)
    if raise_on_failure and info.status != nat.OK:
        nat.check(info.status)
    return x, info


def bicgstab(A: ColoredMatrix, b, x0=None, *, tol: float = 1e-10, maxiter: int | None = None,
             raise_on_failure: bool = True):
    """Solve A x = b for a general (non-symmetric) CSR matrix, FP64,
    Jacobi-preconditioned BiCGSTAB; returns (x, PcgInfo)."""
    t = nat.torch()
    n = A.shape[0]
    bv = nat.to_device(b).reshape(-1).contiguous()
    x = t.zeros(n, dtype=t.float64, device=bv.device) if x0 is None else \
        nat.to_device(x0).reshape(-1).clone()
    L = nat.lib()
    ws = t.empty(int(L.tatva_bicgstab_csr_workspace(n)), dtype=t.uint8, device=bv.device)
    st = nat.PcgStats()
    nat.check(L.tatva_bicgstab_csr(n, A.nnz, nat.dptr(A.indptr), nat.dptr(A.indices),
                                   nat.dptr(A.data), nat.dptr(bv), nat.dptr(x), float(tol),
                                   10 * n if maxiter is None else int(maxiter), nat.dptr(ws),
                                   ws.numel(), ctypes.byref(st), nat.stream()))
    info = _stats_to_info(st)
    if raise_on_failure and info.status != nat.OK:
        nat.check(info.status)
    return x, info


def pcg_matfree(energy: ReducedEnergy, u_full32, b32, x0=None, *, diag32=None, tol: float = 1e-4,
                maxiter: int | None = None, chunk: int = 16, workspace=None):
    """FP32 matrix-free Jacobi-PCG for H(u) x = b on the free dofs; returns (x32, PcgInfo).

    ``u_full32`` is the FP32 linearisation state (full vector), ``b32`` reduced FP32.
    """
    t = nat.torch()
    if not isinstance(energy, ReducedEnergy):
        raise TypeError("pcg_matfree needs energy.reduced(lifter)")
    n = energy.size
    b = nat.to_device(b32, t.float32).reshape(-1).contiguous()
    x = t.zeros(n, dtype=t.float32, device=b.device) if x0 is None else \
        nat.to_device(x0, t.float32).reshape(-1).clone()
    u = nat.to_device(u_full32, t.float32).reshape(-1).contiguous()
    lf = energy.lifter
    if diag32 is None:
        if lf.has_periodic:
            diag32 = lf.reduce_adjoint(energy.energy.hessian_diag(u)).to(t.float32)
        else:
            diag32 = energy.energy.hessian_diag(u, dof_map=lf.dof_map_device(), n_out=n,
                                                dtype=t.float32)
    maxiter = 10 * n if maxiter is None else int(maxiter)
    L = nat.lib()
    op = energy.op
    st = nat.PcgStats()
    desc = energy.energy._desc
    if lf.has_periodic:
        # P^T H P: spread through the lift's read map, sum the rows of each reduced dof
        nbytes = L.tatva_pcg_matfree_mapped_workspace(op.handle, n)
        ws = (workspace or _Workspace()).get(nbytes)
        ptr, lst = lf.reduce_rows_device()
        nat.check(L.tatva_pcg_matfree_f32_mapped(
            op.handle, ctypes.byref(desc), nat.dptr(u), nat.dptr(lf.read_map_device()),
            nat.dptr(ptr), nat.dptr(lst), n, nat.dptr(diag32), nat.dptr(b), nat.dptr(x),
            float(tol), maxiter, int(chunk), nat.dptr(ws), ws.numel(), ctypes.byref(st),
            nat.stream()))
        return x, _stats_to_info(st)
    nbytes = L.tatva_pcg_matfree_workspace(op.handle, ctypes.byref(desc), n)
    ws = (workspace or _Workspace()).get(nbytes)
    nat.check(L.tatva_pcg_matfree_f32(op.handle, ctypes.byref(desc), nat.dptr(u),
                                      nat.dptr(lf.dof_map_device()), n,
                                      nat.dptr(diag32), nat.dptr(b), nat.dptr(x), float(tol),
                                      maxiter, int(chunk), nat.dptr(ws), ws.numel(),
                                      ctypes.byref(st), nat.stream()))
    return x, _stats_to_info(st)


def set_profiling(enable: bool) -> None:
    """Sample one operator launch per captured PCG chunk with HIP events."""
    nat.check(nat.lib().tatva_set_profiling(1 if enable else 0))


OPT_MATFREE_LIN_CACHE = 1
OPT_MATFREE_GEOM_CACHE = 2   # include/tatva_amd.h TATVA_OPT_MATFREE_GEOM_CACHE
OPT_MATFREE_DEFERRED_UPDATE = 3   # include/tatva_amd.h TATVA_OPT_MATFREE_DEFERRED_UPDATE
OPT_MATFREE_FULL_VECTORS = 4   # include/tatva_amd.h TATVA_OPT_MATFREE_FULL_VECTORS


_LIN_CACHE_MODES = {"never": 0, "auto": 1, "always": 2}


def set_linearization_cache(mode: str) -> None:
    """Matrix-free FP32 PCG: linearise the density once per solve and read it in
    every Hessian-vector product ("always"), recompute it in every product
    ("never", the default: faster on this hardware, measured), or cache it on
    Hex8 meshes only ("auto").  The results agree to FP32 rounding."""
    if mode not in _LIN_CACHE_MODES:
        raise ValueError(f"mode must be one of {sorted(_LIN_CACHE_MODES)}")
    nat.check(nat.lib().tatva_set_option(OPT_MATFREE_LIN_CACHE, _LIN_CACHE_MODES[mode]))


def set_geometry_cache(on: bool) -> None:
    """Matrix-free FP32 PCG on affine Hex8 meshes (every element a
    parallelepiped): write each element's Jacobian inverse (FP32) and det J
    (FP64) once per solve and read them in every Hessian-vector product instead
    of the node coordinates (on, the default), or compute them from the
    coordinates in every product (off).  The same result bit for bit; 44 bytes
    per element of workspace."""
    nat.check(nat.lib().tatva_set_option(OPT_MATFREE_GEOM_CACHE, 1 if on else 0))


def set_deferred_update(on: bool) -> None:
    """Matrix-free FP32 PCG: apply each CG iteration's p and x updates in the
    tile kernel of the next Hessian-vector product, which has memory bandwidth
    to spare (on, the default), or in the iteration's own step kernel (off).
    The same iterates bit for bit."""
    nat.check(nat.lib().tatva_set_option(OPT_MATFREE_DEFERRED_UPDATE, 1 if on else 0))


def set_full_vectors(on: bool) -> None:
    """Matrix-free FP32 PCG with Fixed constraints: run the CG on full-length
    vectors (every dof of the mesh, the inverse diagonal 0 on the constrained
    ones) so that no kernel of the iteration reads the dof map (on, the
    default), or on the reduced vectors (off).  The products are the same bits;
    the dot products sum in another order."""
    nat.check(nat.lib().tatva_set_option(OPT_MATFREE_FULL_VECTORS, 1 if on else 0))


def dot(x, y):
    """x . y (FP64 device scalar) with the deterministic reduction kernel; its
    partial sums go to a workspace allocated on the current stream."""
    t = nat.torch()
    L = nat.lib()
    n = x.numel()
    out = t.empty((), dtype=t.float64, device=x.device)
    ws = t.empty(int(L.tatva_dot_workspace(n)), dtype=t.uint8, device=x.device)
    nat.check(L.tatva_dot(n, nat.dptr(x), nat.dptr(y), nat.dptr(out), nat.dptr(ws), ws.numel(),
                          nat.stream()))
    return out


def norm(x) -> float:
    """Euclidean norm with the deterministic FP64 reduction kernel."""
    return math.sqrt(max(float(dot(x, x).item()), 0.0))


@dataclass
class NewtonStats:
    newton_iterations: int = 0
    cg_iterations: int = 0
    residual_history: list = field(default_factory=list)
    converged: bool = False
    seconds: float = 0.0
    cg_seconds: float = 0.0
    assembly_seconds: float = 0.0
    op_ms: float = 0.0
    op_samples: int = 0
    op_main_ms: float = 0.0


def padded_node_layout(mesh, lifter, nc):
    """Host part of the whole-node layout (``_PaddedNodes``): the lifter that
    removes only nodes with no free dof, its pattern S (node-blocked, sorted),
    each reduced dof's index in it (red_to_pad), the positions of S's entries
    in rows or columns of constrained dofs (off the diagonal: zeroed) and of
    their diagonal entries (set to 1).  None when the reduced pattern is
    node-blocked already."""
    from .lifter import Fixed, Lifter  # noqa: PLC0415
    from .sparse import pattern_from_mesh  # noqa: PLC0415

    free = lifter.free_dofs
    nodes = np.unique(free // nc)
    pad_dofs = (nodes[:, None] * nc + np.arange(nc)).ravel()
    if pad_dofs.size == free.size:
        return None
    mask = np.ones(lifter.size, dtype=bool)
    mask[pad_dofs] = False
    L_pad = Lifter(lifter.size, Fixed(np.nonzero(mask)[0], 0.0))
    S = L_pad.adapt_sparsity(pattern_from_mesh(mesh, nc))
    S.sort_indices()
    ip, ix = S.indptr.astype(np.int64), S.indices
    extra = np.nonzero(~np.isin(pad_dofs, free))[0]
    lens = ip[extra + 1] - ip[extra]
    row = np.repeat(extra, lens)                                      # entries (d, i)
    pos = np.repeat(ip[extra] - np.cumsum(np.r_[0, lens[:-1]]), lens) + np.arange(lens.sum())
    col = ix[pos].astype(np.int64)
    on = col == row
    diag_pos = pos[on]
    # the symmetric entry (i, d) of each off-diagonal (d, i): bisection in row i
    i, d = col[~on], row[~on]
    lo, hi = ip[i], ip[i + 1]
    while np.any(lo < hi):
        mid = (lo + hi) // 2
        go = (lo < hi) & (ix[np.minimum(mid, ix.size - 1)] < d)
        lo, hi = np.where(go, mid + 1, lo), np.where(go | (lo >= hi), hi, mid)
    zero_pos = np.unique(np.concatenate([pos[~on], lo]))
    red_to_pad = np.searchsorted(pad_dofs, free)
    return L_pad, S, red_to_pad, zero_pos, diag_pos


class _PaddedNodes:
    """Whole-node layout of the FP64 Newton system under Fixed constraints on
    single node components (e.g. u_x prescribed on a face: those nodes keep
    their other components).  The reduced Hessian S[free][:, free]
    (lifter/base.py:322-331) then has rows of 2 and 3 dofs per node and no node
    blocks.  Here the system is solved on every dof of each node that has a
    free one: the Hessian on that whole-node pattern, with the rows and columns
    of the constrained dofs among them zeroed and a unit diagonal,

        [K_ff  0] [du_f]   [r_f]
        [ 0    I] [du_c] = [ 0 ]

    so du_c = 0 and du_f solves K_ff du_f = r_f, the reduced Newton system.  CG
    on it builds the same Krylov space (the constrained entries of every vector
    stay exactly 0), and the node-block SpMV streams 8 bytes per value plus one
    int32 per 3x3 block instead of the 12 bytes per value of a sliced-ELL copy
    of the reduced matrix."""

    def __init__(self, lifter, cm, red_to_pad, zero_pos, diag_pos):
        self.lifter, self.cm = lifter, cm
        self.red_to_pad, self.zero_pos, self.diag_pos = red_to_pad, zero_pos, diag_pos

    @classmethod
    def build(cls, energy, lifter, nc):
        from .sparse import ColoredMatrix  # noqa: PLC0415

        lay = padded_node_layout(energy.op.mesh, lifter, nc)
        if lay is None:
            return None          # the reduced matrix is node-blocked already
        L_pad, S, red_to_pad, zero_pos, diag_pos = lay
        t = nat.torch()
        cm = ColoredMatrix.from_csr(S)
        return cls(L_pad, cm, nat.to_device(red_to_pad, t.int64),
                   nat.to_device(zero_pos, t.int64), nat.to_device(diag_pos, t.int64))

    def assemble(self, asm, u_full, data):
        data = asm.assemble_full_state(u_full, self.lifter, data)
        data.index_fill_(0, self.zero_pos, 0.0)
        data.index_fill_(0, self.diag_pos, 1.0)
        return data

    def spread(self, r):
        t = nat.torch()
        out = t.zeros(self.cm.shape[0], dtype=r.dtype, device=r.device)
        out.index_copy_(0, self.red_to_pad, r)
        return out

    def gather(self, x):
        return x.index_select(0, self.red_to_pad)


class Newton:
    """Newton's method on the free dofs of ``lifter`` for a built-in energy.

    ``solve(u_full, lifter)`` updates the full FP64 state in place until
    ||r_k|| <= tol * ||r_0|| (r = reduced residual) and returns NewtonStats.
    """

    def __init__(self, energy: Energy, lifter, *, mode: str = "csr", tol: float = 1e-8,
                 lin_tol: float = 1e-10, inner_tol: float = 1e-3, max_newton: int = 50,
                 maxiter: int | None = None, chunk: int = 32, colored_matrix=None,
                 abs_tol: float = 0.0):
        if mode not in ("csr", "mixed"):
            raise ValueError("mode must be 'csr' or 'mixed'")
        self.energy = energy
        self.mode = mode
        self.tol, self.lin_tol, self.inner_tol = tol, lin_tol, inner_tol
        self.abs_tol = abs_tol
        self.max_newton = max_newton
        self.maxiter = maxiter
        self.chunk = chunk
        self.red = energy.reduced(lifter)
        # node-block SpMV for vector fields (falls back to CSR if the pattern is not blocked)
        self.block_size = energy.ncomp if energy.ncomp in (2, 3) else None
        self._cm = colored_matrix
        self._asm = None
        self._ws = _Workspace()
        self._pad = None
        if mode == "csr":
            if colored_matrix is None and self.block_size and not lifter.has_periodic:
                # constraints on single components of a node would break the
                # node blocks of the reduced matrix: solve on whole nodes
                self._pad = _PaddedNodes.build(energy, lifter, self.block_size)
            if self._pad is not None:
                self._cm = self._pad.cm
                self._asm = _HessianAssembler(energy.reduced(self._pad.lifter).grad, self._pad.cm)
            else:
                if colored_matrix is None:
                    from .sparse import ColoredMatrix, pattern_from_mesh  # noqa: PLC0415
                    pat = lifter.adapt_sparsity(pattern_from_mesh(energy.op.mesh, energy.ncomp))
                    colored_matrix = ColoredMatrix.from_csr(pat)
                self._cm = colored_matrix
                self._asm = _HessianAssembler(self.red.grad, colored_matrix)

    @property
    def colored_matrix(self):
        """The matrix the FP64 PCG solves with: the reduced Hessian's pattern, or
        with constrained single components the whole-node pattern (see
        ``_PaddedNodes``)."""
        return self._cm

    def _check_layout(self, lifter) -> None:
        """The reduced sizes, CSR pattern and kernel output lengths were fixed by the
        constructor's lifter: a lifter for ``solve`` may change constraint values
        (``lifter.at[key].set``), not which dofs are constrained or tied."""
        cur = self.red.lifter
        if lifter is cur or lifter._layout is cur._layout:
            return
        if (lifter.size != cur.size or lifter.size_reduced != cur.size_reduced
                or not np.array_equal(lifter.read_map, cur.read_map)
                or not np.array_equal(lifter.dof_map, cur.dof_map)):
            raise ValueError("Newton.solve: the lifter constrains other dofs than the one the "
                             "solver was built with; build a new Newton for it")

    def solve(self, u_full, lifter=None) -> NewtonStats:
        t = nat.torch()
        if lifter is not None:
            self._check_layout(lifter)
            self.red.lifter = lifter
        lf = self.red.lifter
        lf.apply_constraints(u_full)
        dof_map = lf.dof_map_device()
        free = lf.free_device()
        n = self.red.size
        stats = NewtonStats()
        L = nat.lib()
        t0 = time.perf_counter()
        r0 = None
        maxiter = self.maxiter if self.maxiter is not None else 10 * n
        data = None
        for it in range(self.max_newton + 1):
            r = self.red.residual_full_state(u_full)
            rn = norm(r)
            stats.residual_history.append(rn)
            if r0 is None:
                r0 = rn
            if not math.isfinite(rn):
                break
            if rn <= max(self.tol * r0, self.abs_tol) or rn == 0.0:
                stats.converged = True
                break
            if it == self.max_newton:
                break
            stats.newton_iterations += 1
            if self.mode == "csr":
                ta = time.perf_counter()
                pad = self._pad
                if pad is None:
                    data = self._asm.assemble_full_state(u_full, lf, data)
                else:
                    data = pad.assemble(self._asm, u_full, data)
                K = replace(self._cm, data=data)
                t.cuda.synchronize()
                tb = time.perf_counter()
                stats.assembly_seconds += tb - ta
                # K du = r, then u <- u - du (the Newton step -K^-1 r)
                rhs = r if pad is None else pad.spread(r)
                du, info = pcg(K, rhs, tol=self.lin_tol, maxiter=maxiter, chunk=self.chunk,
                               workspace=self._ws, block_size=self.block_size)
                if pad is not None:
                    du = pad.gather(du)
                stats.cg_seconds += time.perf_counter() - tb
                stats.cg_iterations += info.iterations
                stats.op_ms += info.op_ms
                stats.op_samples += info.op_samples
                stats.op_main_ms += info.op_main_ms
                if info.status == nat.EBREAKDOWN:
                    break
                nat.check(L.tatva_axpy_free(n, nat.dptr(free), -1.0, nat.dptr(du),
                                            nat.dptr(u_full), nat.stream()))
                if lf.has_periodic:
                    lf.apply_constraints(u_full)  # slaves follow their masters
            else:
                u32 = t.empty(u_full.numel(), dtype=t.float32, device=u_full.device)
                nat.check(L.tatva_cast_f64_f32(u_full.numel(), 1.0, nat.dptr(u_full), nat.dptr(u32),
                                               nat.stream()))
                b32 = t.empty(n, dtype=t.float32, device=u_full.device)
                nat.check(L.tatva_cast_f64_f32(n, -1.0, nat.dptr(r), nat.dptr(b32), nat.stream()))
                tb = time.perf_counter()
                du, info = pcg_matfree(self.red, u32, b32, tol=self.inner_tol, maxiter=maxiter,
                                       chunk=self.chunk, workspace=self._ws)
                stats.cg_seconds += time.perf_counter() - tb
                stats.cg_iterations += info.iterations
                stats.op_ms += info.op_ms
                stats.op_samples += info.op_samples
                stats.op_main_ms += info.op_main_ms
                if info.status == nat.EBREAKDOWN:
                    break
                nat.check(L.tatva_axpy_free_f32(n, nat.dptr(free), 1.0, nat.dptr(du),
                                                nat.dptr(u_full), nat.stream()))
                if lf.has_periodic:
                    lf.apply_constraints(u_full)  # slaves follow their masters
        t.cuda.synchronize()
        stats.seconds = time.perf_counter() - t0
        return stats


__all__ = ["pcg", "bicgstab", "pcg_matfree", "PcgInfo", "Newton", "NewtonStats", "norm", "set_profiling"]
