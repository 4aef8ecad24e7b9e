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

    @property
    def converged(self) -> bool:
        return self.status == nat.OK


def _stats_to_info(st: nat.PcgStats) -> PcgInfo:
    return PcgInfo(int(st.iterations), float(st.rel_residual), int(st.status), float(st.b_norm),
                   float(st.op_ms), int(st.op_samples))


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
    nbytes = L.tatva_pcg_csr_workspace(n)
    ws = (workspace or _WS_CSR).get(nbytes)
    st = nat.PcgStats()
    view = A.block_view(block_size) if block_size else None
    bc, bs = (view[0], view[1]) if view is not None else (None, 1)
    nat.check(L.tatva_pcg_csr(n, A.nnz, nat.dptr(A.indptr), nat.dptr(A.indices), nat.dptr(A.data),
                              nat.dptr(bc), bs, nat.dptr(bv), nat.dptr(x), float(tol), maxiter,
                              int(chunk), nat.dptr(ws), ws.numel(), ctypes.byref(st),
                              nat.stream()))
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
    nbytes = L.tatva_pcg_matfree_workspace(op.handle, n)
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


def norm(x) -> float:
    """Euclidean norm with the deterministic FP64 reduction kernel."""
    t = nat.torch()
    out = t.empty((), dtype=t.float64, device=x.device)
    nat.check(nat.lib().tatva_dot(x.numel(), nat.dptr(x), nat.dptr(x), nat.dptr(out), nat.stream()))
    return math.sqrt(max(float(out.item()), 0.0))


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
        if mode == "csr":
            if colored_matrix is None:
                from .sparse import ColoredMatrix, pattern_from_mesh  # noqa: PLC0415
                pat = lifter.adapt_sparsity(pattern_from_mesh(energy.op.mesh, energy.ncomp))
                colored_matrix = ColoredMatrix.from_csr(pat)
            self._cm = colored_matrix
            self._asm = _HessianAssembler(self.red.grad, colored_matrix)

    @property
    def colored_matrix(self):
        return self._cm

    def _check_layout(self, lifter) -> None:
        """The reduced sizes, CSR pattern and kernel output lengths were fixed by the
        constructor's lifter: a lifter for ``solve`` may change constraint values
        (``lifter.at[key].set``), not which dofs are constrained or tied."""
        cur = self.red.lifter
        if lifter is cur or lifter._layout is cur._layout:
            return
        import numpy as np  # noqa: PLC0415

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
                data = self._asm.assemble_full_state(u_full, lf, data)
                K = replace(self._cm, data=data)
                t.cuda.synchronize()
                tb = time.perf_counter()
                stats.assembly_seconds += tb - ta
                # K du = r, then u <- u - du (the Newton step -K^-1 r)
                du, info = pcg(K, r, tol=self.lin_tol, maxiter=maxiter, chunk=self.chunk,
                               workspace=self._ws, block_size=self.block_size)
                stats.cg_seconds += time.perf_counter() - tb
                stats.cg_iterations += info.iterations
                stats.op_ms += info.op_ms
                stats.op_samples += info.op_samples
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
                if info.status == nat.EBREAKDOWN:
                    break
                nat.check(L.tatva_axpy_free_f32(n, nat.dptr(free), 1.0, nat.dptr(du),
                                                nat.dptr(u_full), nat.stream()))
                if lf.has_periodic:
                    lf.apply_constraints(u_full)  # slaves follow their masters
        t.cuda.synchronize()
        stats.seconds = time.perf_counter() - t0
        return stats


__all__ = ["pcg", "pcg_matfree", "PcgInfo", "Newton", "NewtonStats", "norm", "set_profiling"]
