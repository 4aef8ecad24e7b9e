"""Lifter: reduced <-> full dof vectors under Dirichlet and periodic constraints.

Mirrors ``tatva.lifter`` (lifter/base.py:55-331, lifter/constraints.py:184-220,
290-318, lifter/common.py:43-65) for ``Fixed`` constraints with constant, array or
``RuntimeValue`` values (load stepping via ``lifter.at[key].set(v)``) and
``Periodic`` constraints.  The
free/constrained index sets are host setup; lift/reduce run as HIP kernels on
GPU tensors.  The reduced sparsity is built natively (S[free][:, free]).
"""

from __future__ import annotations

import ctypes
from collections.abc import Hashable, Mapping
from uuid import uuid4

import numpy as np
import scipy.sparse as sps

from . import _native as nat


class LifterError(ValueError):
    """Problem with the lifter, e.g. a missing runtime value (lifter/common.py:43)."""


class RuntimeValue:
    """A constraint value provided at run time (lifter/common.py:47-65)."""

    def __init__(self, key: Hashable | None = None, default=None):
        self.key = key if key is not None else uuid4()
        self.default = default

    def get_value(self, runtime_values: Mapping):
        if self.key not in runtime_values:
            raise LifterError(f"Runtime value for (key={self.key}) not set on lifter")
        return runtime_values[self.key]


class Constraint:
    """Base class of constraints (lifter/constraints.py:55-181)."""

    dofs: np.ndarray

    def __init__(self, dofs):
        self.dofs = _as_index(dofs)
        self._id = uuid4()

    def __eq__(self, other) -> bool:
        return type(self) is type(other) and self._id == other._id

    def __hash__(self) -> int:
        return hash((type(self), self._id))

    def runtime_specs(self):
        return ()

    def augment_sparsity(self, sparsity):
        return sparsity


def _as_index(dofs) -> np.ndarray:
    try:
        import torch  # noqa: PLC0415
        if isinstance(dofs, torch.Tensor):
            dofs = dofs.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    a = np.asarray(dofs)
    if a.size and a.dtype.kind not in "iu":
        raise LifterError("constraint dofs must be integer indices")
    return a.astype(np.int64).ravel()


class Fixed(Constraint):
    """Dirichlet condition: ``dofs`` are set to ``values`` (lifter/constraints.py:290-318)."""

    def __init__(self, dofs, values=None):
        super().__init__(dofs)
        self.values = 0.0 if values is None else values
        self._dev_cache = None

    def runtime_specs(self):
        return (self.values,) if isinstance(self.values, RuntimeValue) else ()

    def resolved(self, runtime_values):
        v = self.values
        if isinstance(v, RuntimeValue):
            v = v.get_value(runtime_values)
        return v

    def apply_lift(self, u_full, runtime_values) -> None:
        """u_full[dofs] = values (in place, on the GPU)."""
        if self.dofs.size == 0:
            return
        L = nat.lib()
        v = self.resolved(runtime_values)
        idx = self._dev_dofs()
        if np.ndim(v) == 0 and not hasattr(v, "shape"):
            nat.check(L.tatva_set_dofs(self.dofs.size, nat.dptr(idx), None, float(v),
                                       nat.dptr(u_full), nat.stream()))
            return
        vals = nat.to_device(v).reshape(-1)
        if vals.numel() == 1:
            vals = vals.expand(self.dofs.size).contiguous()
        if vals.numel() != self.dofs.size:
            raise LifterError(f"Fixed has {self.dofs.size} dofs but {vals.numel()} values")
        nat.check(L.tatva_set_dofs(self.dofs.size, nat.dptr(idx), nat.dptr(vals.contiguous()), 0.0,
                                   nat.dptr(u_full), nat.stream()))

    def _dev_dofs(self):
        if self._dev_cache is None:
            self._dev_cache = nat.to_device(self.dofs, nat.torch().int32)
        return self._dev_cache


class Periodic(Constraint):
    """``dofs`` take the values of ``master_dofs`` (lifter/constraints.py:184-220)."""

    def __init__(self, dofs, master_dofs):
        super().__init__(dofs)
        self.master_dofs = _as_index(master_dofs)
        if self.master_dofs.size != self.dofs.size:
            raise LifterError("Periodic needs one master per constrained dof")
        self._dev_cache = None

    def augment_sparsity(self, sparsity):
        """M^T S M with M = I + (slave -> master) (constraints.py:194-212); the
        product runs in int32 so that no entry count can wrap."""
        S = sps.csr_matrix(sparsity)
        n = S.shape[0]
        rows = np.concatenate([np.arange(n), self.dofs])
        cols = np.concatenate([np.arange(n), self.master_dofs])
        M = sps.csr_matrix((np.ones(rows.size, dtype=np.int32), (rows, cols)), shape=(n, n))
        S1 = sps.csr_matrix((np.ones(S.nnz, dtype=np.int32), S.indices, S.indptr), shape=S.shape)
        out = (M.T @ (S1 @ M)).tocsr()
        out.sum_duplicates()
        out.sort_indices()
        return sps.csr_matrix((np.ones(out.nnz, dtype=np.int8), out.indices, out.indptr),
                              shape=out.shape)

    def apply_lift(self, u_full, runtime_values) -> None:
        """u_full[dofs] = u_full[master_dofs] (constraints.py:214-216), in place:
        all masters are read before any slave is written."""
        if self.dofs.size == 0:
            return
        t = nat.torch()
        L = nat.lib()
        slaves, masters = self._dev()
        tmp = t.empty(self.dofs.size, dtype=t.float64, device=u_full.device)
        nat.check(L.tatva_gather_free(self.dofs.size, nat.dptr(masters), nat.dptr(u_full),
                                      nat.dptr(tmp), nat.stream()))
        nat.check(L.tatva_scatter_free(self.dofs.size, nat.dptr(slaves), nat.dptr(tmp),
                                       nat.dptr(u_full), nat.stream()))

    def _dev(self):
        if self._dev_cache is None:
            i32 = nat.torch().int32
            self._dev_cache = (nat.to_device(self.dofs, i32), nat.to_device(self.master_dofs, i32))
        return self._dev_cache


class _Layout:
    """Free/constrained index sets shared by lifters that differ in runtime values.

    ``dof_map``: full dof -> reduced index of a free dof, else -1.
    ``read_map``: the linear part of the lift, full dof -> reduced index whose
    value the dof takes (a free dof its own, a Periodic slave its master's,
    resolved in constraint order as lifter/base.py:201-221 applies them), else -1.
    ``red_ptr``/``red_list``: for each reduced index the full dofs reading it,
    ascending; reduce_adjoint sums them (lifter/base.py:235-251).
    """

    def __init__(self, size: int, constraints):
        self.size = size
        if constraints:
            con = np.unique(np.concatenate([c.dofs for c in constraints]))
        else:
            con = np.zeros(0, dtype=np.int64)
        if con.size and (con[0] < 0 or con[-1] >= size):
            raise LifterError("constrained dofs outside [0, size)")
        for c in constraints:
            if isinstance(c, Periodic) and c.master_dofs.size and \
                    (c.master_dofs.min() < 0 or c.master_dofs.max() >= size):
                raise LifterError("periodic master dofs outside [0, size)")
        self.constrained = con
        mask = np.ones(size, dtype=bool)
        mask[con] = False
        self.free = np.nonzero(mask)[0].astype(np.int64)
        self.dof_map = np.full(size, -1, dtype=np.int32)
        self.dof_map[self.free] = np.arange(self.free.size, dtype=np.int32)
        self.has_periodic = any(isinstance(c, Periodic) for c in constraints)
        rm = self.dof_map.copy()
        for c in constraints:
            if isinstance(c, Periodic):
                rm[c.dofs] = rm[c.master_dofs].copy()
            else:
                rm[c.dofs] = -1
        self.read_map = rm
        order = np.argsort(rm, kind="stable")
        order = order[rm[order] >= 0]
        counts = np.bincount(rm[order], minlength=self.free.size)
        self.red_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        self.red_list = order.astype(np.int32)
        self._free_dev = None
        self._map_dev = None
        self._read_dev = None
        self._rows_dev = None

    def free_dev(self):
        if self._free_dev is None:
            self._free_dev = nat.to_device(self.free, nat.torch().int32)
        return self._free_dev

    def map_dev(self):
        if self._map_dev is None:
            self._map_dev = nat.to_device(self.dof_map, nat.torch().int32)
        return self._map_dev

    def read_dev(self):
        if self._read_dev is None:
            self._read_dev = nat.to_device(self.read_map, nat.torch().int32)
        return self._read_dev

    def rows_dev(self):
        if self._rows_dev is None:
            i32 = nat.torch().int32
            self._rows_dev = (nat.to_device(self.red_ptr, i32), nat.to_device(self.red_list, i32))
        return self._rows_dev


class Lifter:
    """``Lifter(size, *constraints)`` (lifter/base.py:55-331)."""

    def __init__(self, size: int, /, *constraints: Constraint):
        self.size = int(size)
        self.constraints = tuple(constraints)
        specs = [s for c in self.constraints for s in c.runtime_specs()]
        self._runtime_keys = tuple(s.key for s in specs)
        self._runtime_values = {s.key: s.default for s in specs if s.default is not None}
        self._layout = _Layout(self.size, self.constraints)

    # ---------------------------------------------------------- properties --

    @property
    def free_dofs(self) -> np.ndarray:
        return self._layout.free

    @property
    def constrained_dofs(self) -> np.ndarray:
        return self._layout.constrained

    @property
    def size_reduced(self) -> int:
        return int(self._layout.free.size)

    @property
    def dof_map(self) -> np.ndarray:
        """full dof -> reduced index, -1 on constrained dofs."""
        return self._layout.dof_map

    def dof_map_device(self):
        return self._layout.map_dev()

    def free_device(self):
        return self._layout.free_dev()

    @property
    def has_periodic(self) -> bool:
        return self._layout.has_periodic

    @property
    def read_map(self) -> np.ndarray:
        """full dof -> reduced index it reads in a lift (-1: constraint value)."""
        return self._layout.read_map

    def read_map_device(self):
        return self._layout.read_dev()

    def reduce_rows_device(self):
        """(ptr, list): the full dofs summed into each reduced index by reduce_adjoint."""
        return self._layout.rows_dev()

    def _require_device_kernels(self):
        for c in self.constraints:
            if not isinstance(c, (Fixed, Periodic)):
                raise NotImplementedError(f"{type(c).__name__} constraints are not supported on "
                                          "the GPU path (Fixed and Periodic)")

    def __hash__(self) -> int:
        return hash((self.size, self.constraints))

    def __eq__(self, other) -> bool:
        if not isinstance(other, Lifter) or self.size != other.size:
            return False
        if self.constraints != other.constraints:
            return False
        a, b = self._runtime_values, other._runtime_values
        if a.keys() != b.keys():
            return False
        return all(np.array_equal(np.asarray(_host(a[k])), np.asarray(_host(b[k]))) for k in a)

    # ----------------------------------------------------- runtime values --

    @property
    def at(self) -> "_RuntimeValueIndexer":
        return _RuntimeValueIndexer(self)

    def with_values(self, updates: Mapping) -> "Lifter":
        for key in updates:
            if key not in self._runtime_keys:
                raise LifterError(
                    f"There is no runtime value with key={key} in the lifter's constraints")
        new = object.__new__(Lifter)
        new.__dict__ = dict(self.__dict__)
        new._runtime_values = {**self._runtime_values, **updates}
        return new

    def add(self, condition: Constraint) -> "Lifter":
        return Lifter(self.size, *self.constraints, condition)

    # --------------------------------------------------------- lift/reduce --

    def apply_constraints(self, u_full) -> None:
        """Write the constrained values into a full GPU vector, in place."""
        self._require_device_kernels()
        for c in self.constraints:
            c.apply_lift(u_full, self._runtime_values)

    def lift(self, u_reduced, u_full):
        """Full vector with free dofs = u_reduced and constraints applied (base.py:201-221)."""
        self._require_device_kernels()
        out = nat.to_device(u_full).reshape(-1).clone()
        if out.shape[0] != self.size:
            raise LifterError(f"u_full must have {self.size} entries")
        self._scatter(u_reduced, out)
        self.apply_constraints(out)
        return out

    def lift_from_zeros(self, u_reduced):
        """``lift`` starting from zeros (base.py:223-229)."""
        self._require_device_kernels()
        t = nat.torch()
        out = t.zeros(self.size, dtype=t.float64, device=nat.device())
        self._scatter(u_reduced, out)
        self.apply_constraints(out)
        return out

    def _scatter(self, u_reduced, out):
        ur = nat.to_device(u_reduced).reshape(-1).contiguous()
        if ur.shape[0] != self.size_reduced:
            raise LifterError(f"u_reduced must have {self.size_reduced} entries, got {ur.shape[0]}")
        nat.check(nat.lib().tatva_scatter_free(self.size_reduced, nat.dptr(self.free_device()),
                                               nat.dptr(ur), nat.dptr(out), nat.stream()))

    def reduce(self, u_full):
        """u_full[free_dofs] (base.py:231-233)."""
        t = nat.torch()
        uf = nat.to_device(u_full).reshape(-1).contiguous()
        if uf.shape[0] != self.size:
            raise LifterError(f"u_full must have {self.size} entries")
        out = t.empty(self.size_reduced, dtype=t.float64, device=uf.device)
        nat.check(nat.lib().tatva_gather_free(self.size_reduced, nat.dptr(self.free_device()),
                                              nat.dptr(uf), nat.dptr(out), nat.stream()))
        return out

    def reduce_adjoint(self, r_full):
        """Reduced dual vector (base.py:235-251): the transposes of the constraints
        in reverse order, then the free rows.  With Fixed only this is
        r_full[free_dofs]; a Periodic slave's entry is added to its master's."""
        self._require_device_kernels()
        if not self.has_periodic:
            return self.reduce(r_full)
        t = nat.torch()
        rf = nat.to_device(r_full).reshape(-1).contiguous()
        if rf.shape[0] != self.size:
            raise LifterError(f"r_full must have {self.size} entries")
        out = t.empty(self.size_reduced, dtype=t.float64, device=rf.device)
        ptr, lst = self.reduce_rows_device()
        nat.check(nat.lib().tatva_reduce_rows(self.size_reduced, nat.dptr(ptr), nat.dptr(lst),
                                              nat.dptr(rf), nat.dptr(out), nat.stream()))
        return out

    def lift_linear(self, v_reduced):
        """Linear part of the lift (constraint values at zero): the tangent of
        lift_from_zeros, v_full[d] = v_reduced[read_map[d]] or 0."""
        t = nat.torch()
        vr = nat.to_device(v_reduced).reshape(-1).contiguous()
        if vr.shape[0] != self.size_reduced:
            raise LifterError(f"v_reduced must have {self.size_reduced} entries")
        out = t.empty(self.size, dtype=t.float64, device=vr.device)
        nat.check(nat.lib().tatva_lift_map(self.size, nat.dptr(self.read_map_device()),
                                           nat.dptr(vr), nat.dptr(out), nat.stream()))
        return out

    # ------------------------------------------------------------ sparsity --

    def adapt_sparsity(self, sparsity):
        """Augment (no-op for Fixed) and reduce to the free dofs (base.py:281-296)."""
        return self.reduce_sparsity(self.augment_sparsity(sparsity))

    def augment_sparsity(self, sparsity):
        """base.py:298-320: pad to ``size`` and let each constraint add couplings."""
        S = sps.csr_matrix(sparsity)
        if S.shape[0] < self.size:
            extra = np.full(self.size - S.shape[0], S.indptr[-1], dtype=S.indptr.dtype)
            S = sps.csr_matrix((S.data, S.indices, np.concatenate([S.indptr, extra])),
                               shape=(self.size, self.size))
        for c in self.constraints:
            S = c.augment_sparsity(S)
        return S

    def reduce_sparsity(self, sparsity):
        """S[free][:, free] (base.py:322-331), built natively."""
        S = sps.csr_matrix(sparsity)
        S.sort_indices()
        n = S.shape[0]
        if n != self.size:
            raise LifterError(f"sparsity has {n} rows, lifter size is {self.size}")
        indptr = np.ascontiguousarray(S.indptr, dtype=np.int32)
        indices = np.ascontiguousarray(S.indices, dtype=np.int32)
        h = ctypes.c_void_p()
        L = nat.lib()
        nat.check(L.tatva_csr_submatrix(n, indptr.ctypes.data, indices.ctypes.data,
                                        self.dof_map.ctypes.data, self.size_reduced,
                                        ctypes.byref(h)))
        try:
            nnz = L.tatva_csr_host_nnz(h)
            ip = np.empty(self.size_reduced + 1, dtype=np.int32)
            ix = np.empty(nnz, dtype=np.int32)
            nat.check(L.tatva_csr_host_copy(h, ip.ctypes.data, ix.ctypes.data))
        finally:
            L.tatva_csr_host_free(h)
        m = self.size_reduced
        return sps.csr_matrix((np.ones(nnz, dtype=np.int8), ix, ip), shape=(m, m))


def _host(v):
    try:
        import torch  # noqa: PLC0415
        if isinstance(v, torch.Tensor):
            return v.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return v


class _RuntimeValueSetter:
    def __init__(self, lifter: Lifter, key):
        self.lifter, self.key = lifter, key

    def set(self, value) -> Lifter:
        return self.lifter.with_values({self.key: value})


class _RuntimeValueIndexer:
    """``lifter.at[key].set(v)`` (lifter/base.py:512-537)."""

    def __init__(self, lifter: Lifter):
        self.lifter = lifter

    def __getitem__(self, key) -> _RuntimeValueSetter:
        return _RuntimeValueSetter(self.lifter, key)

    def __call__(self, key) -> _RuntimeValueSetter:
        return _RuntimeValueSetter(self.lifter, key)


def lifted(fn=None, *, argnums=0, output=None):
    """Lift a function of full vectors to one of reduced vectors (lifter/base.py:428-509)."""
    argnums = (argnums,) if isinstance(argnums, int) else tuple(argnums)
    if fn is None:
        return lambda f: lifted(f, argnums=argnums, output=output)

    def lifted_fn(lifter: Lifter, *args, **kwargs):
        args = list(args)
        for i in argnums:
            a = args[i]
            shape = tuple(getattr(a, "shape", ()))
            if shape != (lifter.size_reduced,):
                raise LifterError(f"Argument {i} has shape {shape} but expected "
                                  f"{(lifter.size_reduced,)} for lifting")
            args[i] = lifter.lift_from_zeros(a)
        out = fn(*args, **kwargs)
        if output is None:
            return out
        if output == "primal":
            return lifter.reduce(out)
        if output == "dual":
            return lifter.reduce_adjoint(out)
        raise LifterError(f"Invalid value for output: {output}")

    return lifted_fn


__all__ = ["Lifter", "Fixed", "Periodic", "Constraint", "RuntimeValue", "LifterError", "lifted"]
