"""Colored sparse Jacobians (tatva.sparse: sparse/base.py, sparse/_extraction.py).

* ``pattern_from_mesh(mesh, n_dofs_per_node)`` -> SciPy CSR pattern (int8 ones),
  built natively from the node graph; identical to tatva's sorted, deduplicated
  element-pair pattern (sparse/_extraction.py:37-102).
* ``ColoredMatrix.from_csr(csr, colors=None)`` colors the pattern with the
  natural-order greedy distance-2 coloring (sparse/_coloring.py:136-153) and
  keeps ``data, indptr, indices, colors`` resident on the GPU.
* ``jacfwd(fn, cm, color_batch_size=None)(u)`` / ``linearized_jacfwd`` fill
  ``data`` with the colored Hessian of a built-in energy: per-color JVPs
  (element JVPs with the color seeds) compressed per row in LDS and
  decompressed into CSR by ``data[k] = J[row(k), colors[indices[k]]]``
  (sparse/base.py:139-227,230-270).  ``fn`` must be the residual of a built-in
  energy (``E.grad``); other callables raise.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field, replace

import numpy as np
import scipy.sparse as sps

from . import _native as nat
from .energies import ReducedEnergy, Residual


def pattern_from_mesh(mesh, n_dofs_per_node: int):
    """CSR sparsity of the element connectivity (sparse/_extraction.py:90-102)."""
    L = nat.lib()
    el = np.ascontiguousarray(mesh.elements, dtype=np.int32)
    n_nodes = mesh.coords.shape[0]
    n = n_nodes * int(n_dofs_per_node)
    h = ctypes.c_void_p()
    nat.check(L.tatva_pattern_from_mesh(el.ctypes.data, el.shape[0], el.shape[1], n_nodes,
                                        int(n_dofs_per_node), None, n, ctypes.byref(h)))
    try:
        nnz = L.tatva_csr_host_nnz(h)
        ip = np.empty(n + 1, dtype=np.int32)
        ix = np.empty(nnz, dtype=np.int32)
        nat.check(L.tatva_csr_host_copy(h, ip.ctypes.data, ix.ctypes.data))
    finally:
        L.tatva_csr_host_free(h)
    return sps.csr_matrix((np.ones(nnz, dtype=np.int8), ix, ip), shape=(n, n))


def _host_csr(L, h, n):
    try:
        nnz = L.tatva_csr_host_nnz(h)
        ip = np.empty(n + 1, dtype=np.int32)
        ix = np.empty(nnz, dtype=np.int32)
        nat.check(L.tatva_csr_host_copy(h, ip.ctypes.data, ix.ctypes.data))
    finally:
        L.tatva_csr_host_free(h)
    return sps.csr_matrix((np.ones(nnz, dtype=np.int8), ix, ip), shape=(n, n))


def pattern_from_element_dofs(element_dofs, size: int, diagonal_dofs=None):
    """CSR sparsity of a general element dof table (E, k) with -1 for "no dof":
    all pairs of valid dofs of each element, plus (d, d) for ``diagonal_dofs``.
    The pattern builder behind pattern_from_compound (sparse/_extraction.py:198-238)."""
    ed = np.ascontiguousarray(element_dofs, dtype=np.int32)
    if ed.ndim != 2:
        raise ValueError("element_dofs must be a 2-D (n_elements, dofs_per_element) array")
    dg = np.ascontiguousarray(np.zeros(0) if diagonal_dofs is None else diagonal_dofs,
                              dtype=np.int32).ravel()
    L = nat.lib()
    h = ctypes.c_void_p()
    nat.check(L.tatva_pattern_from_element_dofs(ed.ctypes.data if ed.size else None, ed.shape[0],
                                                ed.shape[1], int(size),
                                                dg.ctypes.data if dg.size else None, dg.size,
                                                ctypes.byref(h)))
    return _host_csr(L, h, int(size))


def pattern_from_fields(mesh, size: int, nodal=(), diagonal=(), block_slices=None):
    """``pattern_from_compound`` (sparse/_extraction.py:125-245) for a state laid
    out as fields, without the Compound class: ``nodal`` is a sequence of
    (n_nodes, dofs_per_node) int tables (global dof of each node's components,
    -1 where a node has none), coupled through the mesh elements; ``diagonal``
    is a sequence of dof arrays that only couple to themselves.  With
    ``block_slices`` (a list of slices, one per block) the blocks
    ``S[s1, s2]`` are returned as a list of lists, as with block_wise=True."""
    el = np.asarray(mesh.elements, dtype=np.int64)
    tables = []
    for nd in nodal:
        nd = np.asarray(nd, dtype=np.int64)
        if nd.ndim == 1:
            nd = nd[:, None]
        if nd.shape[0] != mesh.coords.shape[0]:
            raise ValueError("a nodal field table needs one row per mesh node")
        tables.append(nd[el].reshape(el.shape[0], -1))
    ed = np.concatenate(tables, axis=1) if tables else np.zeros((0, 0), dtype=np.int64)
    dg = np.concatenate([np.asarray(d, dtype=np.int64).ravel() for d in diagonal]) \
        if len(diagonal) else None
    S = pattern_from_element_dofs(ed, size, dg)
    if block_slices is None:
        return S
    return [[S[s1, s2] for s2 in block_slices] for s1 in block_slices]


def distance2_colors(indptr, indices, n_dofs: int) -> np.ndarray:
    """Natural-order greedy distance-2 coloring (sparse/_coloring.py:270-283)."""
    ip = np.ascontiguousarray(indptr, dtype=np.int32)
    ix = np.ascontiguousarray(indices, dtype=np.int32)
    out = np.empty(int(n_dofs), dtype=np.int32)
    nc = ctypes.c_int32()
    nat.check(nat.lib().tatva_distance2_colors(int(n_dofs), ip.ctypes.data, ix.ctypes.data,
                                               out.ctypes.data, ctypes.byref(nc)))
    return out


@dataclass(frozen=True, eq=False)
class ColoredMatrix:
    """CSR matrix with a distance-2 coloring of its columns (sparse/base.py:37-105).

    ``data`` (float64), ``indptr``/``indices``/``colors`` (int32) are GPU tensors.
    """

    data: object = field(repr=False)
    indptr: object = field(repr=False)
    indices: object = field(repr=False)
    shape: tuple
    colors: object = field(repr=False)
    n_colors: int = 0
    _host_indptr: np.ndarray | None = field(default=None, repr=False)
    _host_indices: np.ndarray | None = field(default=None, repr=False)
    # pattern-derived data (block views), shared by copies made with replace()
    _cache: dict = field(default_factory=dict, repr=False)

    @classmethod
    def from_csr(cls, csr_matrix, colors=None) -> "ColoredMatrix":
        S = sps.csr_matrix(csr_matrix)
        S.sort_indices()
        n = S.shape[0]
        if S.shape[0] != S.shape[1]:
            raise ValueError("ColoredMatrix needs a square pattern")
        ip = np.ascontiguousarray(S.indptr, dtype=np.int32)
        ix = np.ascontiguousarray(S.indices, dtype=np.int32)
        if colors is None:
            colors = distance2_colors(ip, ix, n)
        else:
            colors = np.ascontiguousarray(nat_host(colors), dtype=np.int32).ravel()
            if colors.shape[0] != n:
                raise ValueError(f"colors must have {n} entries")
            if colors.size and colors.min() < 0:
                raise ValueError("colors must be non-negative")
            # the decompression data[k] = J_c[row, colors[indices[k]]] is exact only
            # when no row has two columns of one color (a distance-2 coloring)
            bad = ctypes.c_int64(-1)
            nat.check(nat.lib().tatva_check_distance2_colors(
                n, ip.ctypes.data, ix.ctypes.data, colors.ctypes.data, ctypes.byref(bad)))
            if bad.value >= 0:
                raise ValueError(f"colors are not a distance-2 coloring of the pattern: row "
                                 f"{bad.value} has two columns of one color")
        t = nat.torch()
        return cls(
            data=nat.to_device(np.asarray(S.data, dtype=np.float64)),
            indptr=nat.to_device(ip, t.int32),
            indices=nat.to_device(ix, t.int32),
            shape=(int(n), int(n)),
            colors=nat.to_device(colors, t.int32),
            n_colors=int(colors.max()) + 1 if colors.size else 0,
            _host_indptr=ip,
            _host_indices=ix,
        )

    @property
    def nnz(self) -> int:
        return int(self._host_indices.shape[0]) if self._host_indices is not None \
            else int(self.indices.shape[0])

    def host_indptr(self) -> np.ndarray:
        """indptr on the host (kept from construction, else copied once)."""
        if self._host_indptr is None:
            object.__setattr__(self, "_host_indptr", nat_host(self.indptr).astype(np.int32))
        return self._host_indptr

    def to_csr(self):
        """SciPy CSR copy on the host (sparse/base.py:76-81)."""
        return sps.csr_matrix((self.data.detach().cpu().numpy(), self._host_indices,
                               self._host_indptr), shape=self.shape)

    def to_dense(self):
        """Dense host copy (sparse/base.py:102-105)."""
        return self.to_csr().toarray()

    def block_view(self, block_size: int):
        """Node-block view of the pattern: (block_cols tensor, block_size), or None
        when rows do not come in groups sharing aligned block columns.  Cached."""
        cache = self._cache
        if block_size in cache:
            return cache[block_size]
        view = None
        n = self.shape[0]
        if block_size in (2, 3) and n % block_size == 0 and self.nnz % (block_size ** 2) == 0:
            t = nat.torch()
            bc = t.empty(self.nnz // block_size ** 2, dtype=t.int32, device=self.indices.device)
            ok = ctypes.c_int32()
            nat.check(nat.lib().tatva_csr_block_view(n, self.nnz, nat.dptr(self.indptr),
                                                     nat.dptr(self.indices), block_size,
                                                     nat.dptr(bc), ctypes.byref(ok), nat.stream()))
            if ok.value:
                view = (bc, block_size)
        cache[block_size] = view
        return view

    def matvec(self, x, *, block_size: int | None = None):
        """y = A x on the GPU (CSR SpMV, or the node-block view when block_size is given)."""
        t = nat.torch()
        xv = nat.to_device(x).reshape(-1).contiguous()
        y = t.empty(self.shape[0], dtype=t.float64, device=xv.device)
        view = self.block_view(block_size) if block_size else None
        if view is not None:
            nat.check(nat.lib().tatva_spmv_bcsr(self.shape[0], nat.dptr(self.indptr),
                                                nat.dptr(view[0]), view[1], nat.dptr(self.data),
                                                nat.dptr(xv), nat.dptr(y), nat.stream()))
            return y
        nat.check(nat.lib().tatva_spmv_csr(self.shape[0], self.nnz, nat.dptr(self.indptr),
                                           nat.dptr(self.indices), nat.dptr(self.data),
                                           nat.dptr(xv), nat.dptr(y), nat.stream()))
        return y

    def __matmul__(self, x):
        return self.matvec(x)


def nat_host(a):
    try:
        import torch  # noqa: PLC0415
        if isinstance(a, torch.Tensor):
            return a.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return np.asarray(a)


def compute_rows_cols(colored_matrix: ColoredMatrix):
    """(rows, col_colors) per non-zero (sparse/base.py:108-136), on the GPU."""
    t = nat.torch()
    cm = colored_matrix
    rows = t.empty(cm.nnz, dtype=t.int32, device=cm.indices.device)
    cc = t.empty(cm.nnz, dtype=t.int32, device=cm.indices.device)
    nat.check(nat.lib().tatva_rows_cols(cm.shape[0], nat.dptr(cm.indptr), nat.dptr(cm.indices),
                                        nat.dptr(cm.colors), nat.dptr(rows), nat.dptr(cc),
                                        nat.stream()))
    return rows, cc


class _HessianAssembler:
    """Holds the per-mesh workspace of the colored-Hessian kernels."""

    def __init__(self, fn, cm: ColoredMatrix):
        if not isinstance(fn, Residual):
            raise NotImplementedError(
                "sparse.jacfwd differentiates the residual of a built-in energy (E.grad); "
                "arbitrary callables cannot run on the GPU path")
        self.res = fn
        self.cm = cm
        en = fn.energy
        self.reduced = isinstance(en, ReducedEnergy)
        self.full = en.energy if self.reduced else en
        n = en.size
        if cm.shape != (n, n):
            raise ValueError(f"pattern shape {cm.shape} does not match the {n} dofs of the energy")
        if cm.n_colors > 2048:
            raise ValueError("more than 2048 colors are not supported")
        self._ws = None
        self._serial = {}

    def workspace(self):
        if self._ws is None:
            t = nat.torch()
            nbytes = nat.lib().tatva_colored_hessian_workspace(self.full.op.handle, self.full.ncomp)
            self._ws = t.empty(int(nbytes), dtype=t.uint8, device=nat.device())
        return self._ws

    def __call__(self, u, data=None):
        en = self.res.energy
        if self.reduced:
            return self.assemble_full_state(en.lift(u), en.lifter, data)
        return self.assemble_full_state(self.full._vec(u), None, data)

    def _serial_flag(self, lifter) -> int:
        """1 when some element holds two dofs that read the same reduced index
        (a periodic slave and its master in one element)."""
        key = id(lifter._layout)
        if self._serial.get(key) is None:
            nc = self.full.ncomp
            el = self.full.op.mesh.elements.astype(np.int64)
            dofs = (el[:, :, None] * nc + np.arange(nc)).reshape(el.shape[0], -1)
            rm = lifter.read_map[dofs]
            rs = np.sort(np.where(rm >= 0, rm, -1 - np.arange(rm.shape[1])[None, :]), axis=1)
            self._serial[key] = int(np.any(rs[:, 1:] == rs[:, :-1]))
        return self._serial[key]

    def assemble_full_state(self, u_full, lifter=None, data=None):
        """Hessian data on the pattern at a lifted state: reduced by ``lifter``
        (None: the full dofs)."""
        t = nat.torch()
        cm = self.cm
        if data is None:
            data = t.empty(cm.nnz, dtype=t.float64, device=u_full.device)
        ws = self.workspace()
        desc = self.full._desc
        L = nat.lib()
        if lifter is not None and lifter.has_periodic:
            ptr, lst = lifter.reduce_rows_device()
            nat.check(L.tatva_colored_hessian_mapped(
                self.full.op.handle, ctypes.byref(desc), nat.dptr(u_full),
                nat.dptr(lifter.read_map_device()), nat.dptr(ptr), nat.dptr(lst), cm.shape[0],
                nat.dptr(cm.colors), cm.n_colors, nat.dptr(cm.indptr), nat.dptr(cm.indices),
                cm.nnz, self._serial_flag(lifter), nat.dptr(data), nat.dptr(ws), ws.numel(),
                nat.stream()))
            return data
        dof_map = lifter.dof_map_device() if lifter is not None else None
        r2f = lifter.free_device() if lifter is not None else None
        nat.check(L.tatva_colored_hessian(
            self.full.op.handle, ctypes.byref(desc), nat.dptr(u_full), nat.dptr(dof_map),
            nat.dptr(r2f), cm.shape[0], nat.dptr(cm.colors), cm.n_colors, nat.dptr(cm.indptr),
            nat.dptr(cm.indices), cm.nnz, nat.dptr(data), nat.dptr(ws), ws.numel(), nat.stream()))
        return data


def jacfwd(fn, colored_matrix: ColoredMatrix, *, color_batch_size: int | None = None):
    """``sparse.jacfwd`` (sparse/base.py:139-176): returns ``u -> ColoredMatrix``.

    ``color_batch_size`` is accepted for API compatibility: each row's
    compressed Jacobian lives in LDS, so all colors are processed in one pass.
    """
    if color_batch_size is not None and color_batch_size < 1:
        raise ValueError("color_batch_size must be positive")
    asm = _HessianAssembler(fn, colored_matrix)

    def _wrapped(u) -> ColoredMatrix:
        return replace(colored_matrix, data=asm(u))

    _wrapped.assembler = asm
    return _wrapped


def linearized_jacfwd(fn, colored_matrix: ColoredMatrix, *, color_batch_size: int | None = None):
    """``sparse.linearized_jacfwd`` (sparse/base.py:179-227): ``u -> (fn(u), ColoredMatrix)``."""
    if color_batch_size is not None and color_batch_size < 1:
        raise ValueError("color_batch_size must be positive")
    asm = _HessianAssembler(fn, colored_matrix)

    def _wrapped(u):
        return fn(u), replace(colored_matrix, data=asm(u))

    _wrapped.assembler = asm
    return _wrapped


def __getattr__(name):
    deprecated = ("create_sparsity_pattern", "reduce_sparsity_pattern",
                  "create_sparsity_pattern_KKT", "create_sparsity_pattern_master_slave",
                  "get_bc_indices")
    if name in deprecated:
        raise ImportError(f"The function {name!r} is deprecated and does not exist anymore. "
                          "Please use the new sparsity functions.")
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")


__all__ = ["ColoredMatrix", "compute_rows_cols", "distance2_colors", "jacfwd",
           "linearized_jacfwd", "pattern_from_mesh", "pattern_from_element_dofs",
           "pattern_from_fields"]
