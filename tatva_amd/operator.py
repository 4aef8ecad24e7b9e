"""Operator: gather -> per-element / per-quadrature-point map -> weighted reduce.

Mirrors ``tatva.Operator`` (tatva/operator.py:77-397) on gfx950 kernels: the
mesh is uploaded once (coordinates, connectivity, node->element incidence),
and ``eval``/``grad``/``integrate``/``get_integration_weights`` run as HIP
kernels through the C ABI.  Inputs may be numpy arrays or torch tensors;
outputs are FP64 torch tensors on the GPU.

The reference's ``map``/``map_over_elements`` accept arbitrary JAX callables.
A HIP kernel cannot execute those, so here they accept the built-in element
functions (gradient, interpolation, Jacobian determinant) and the built-in
energy densities; any other callable raises ``NotImplementedError`` rather
than running on the host.
"""

from __future__ import annotations

import ctypes
import weakref
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _native as nat
from .element import Element
from .mesh import Mesh


class _DeviceMesh:
    """Owns the native ``tatva_mesh`` handle; freed with the Operator."""

    def __init__(self, mesh: Mesh, element: Element):
        L = nat.lib()
        dev = nat.device()
        h = ctypes.c_void_p()
        coords = np.ascontiguousarray(mesh.coords, dtype=np.float64)
        elems = np.ascontiguousarray(mesh.elements, dtype=np.int32)
        qp = np.ascontiguousarray(element.quad_points, dtype=np.float64)
        qw = np.ascontiguousarray(element.quad_weights, dtype=np.float64)
        nat.check(L.tatva_mesh_create(dev.index, element.kind, coords.ctypes.data, coords.shape[0],
                                      coords.shape[1], elems.ctypes.data, elems.shape[0],
                                      elems.shape[1], qw.shape[0], qp.ctypes.data, qw.ctypes.data,
                                      ctypes.byref(h)))
        self.handle = h.value
        self.device = dev
        self._fin = weakref.finalize(self, L.tatva_mesh_destroy, h.value)

    def close(self):
        self._fin()


def _validate(mesh: Mesh, element: Element) -> None:
    """``Operator.__check_init__`` (tatva/operator.py:132-170)."""
    coords, elements = mesh.coords, mesh.elements
    if coords.ndim != 2:
        raise ValueError("Mesh coordinates must be a 2D array shaped (n_nodes, n_dim).")
    if coords.shape[0] == 0:
        raise ValueError("Mesh must contain at least one node.")
    if elements.ndim != 2:
        raise ValueError("Mesh elements must be a 2D array shaped (n_elements, n_nodes_per_element).")
    if elements.shape[0] == 0:
        raise ValueError("Mesh must contain at least one element.")
    if not np.issubdtype(elements.dtype, np.integer):
        raise TypeError("Mesh element connectivity must contain integer indices.")
    if elements.min() < 0:
        raise ValueError("Mesh element connectivity contains negative node indices.")
    if elements.max() >= coords.shape[0]:
        raise ValueError("Mesh element connectivity references nodes outside the mesh coordinates array.")
    if coords.shape[1] != element.dim:
        raise ValueError(f"{type(element).__name__} needs {element.dim}-D coordinates.")
    if elements.shape[1] != element.nne:
        raise ValueError(f"{type(element).__name__} has {element.nne} nodes per element.")


@dataclass(frozen=True, eq=False)
class Operator:
    """``Operator(mesh, element, batch_size=None, cache_weights=False)``.

    ``batch_size`` is accepted for API compatibility; the kernels stream all
    elements in one launch.  ``cache_weights`` keeps detJ*w resident after the
    first use (operator.py:119-130).
    """

    mesh: Mesh
    element: Element
    batch_size: int | None = None
    cache_weights: bool = False
    _dev: _DeviceMesh | None = field(default=None, init=False, repr=False)
    _w: object = field(default=None, init=False, repr=False)

    def __post_init__(self):
        _validate(self.mesh, self.element)
        if self.batch_size is None:
            object.__setattr__(self, "batch_size", self.mesh.n_elements)

    # ------------------------------------------------------------ plumbing --

    @property
    def handle(self) -> int:
        if self._dev is None:
            object.__setattr__(self, "_dev", _DeviceMesh(self.mesh, self.element))
        return self._dev.handle

    def device_layout(self) -> dict:
        """Sizes of the mesh's device layout (tiles of elements and their
        (tile, local node) slots), for the kernels' byte counts."""
        out = (ctypes.c_int64 * 8)()
        nat.check(nat.lib().tatva_mesh_info(self.handle, out, 8))
        keys = ("n_nodes", "n_elements", "n_tiles", "n_slots", "max_local", "max_colors",
                "hex_affine", "n_quad")
        return {k: int(v) for k, v in zip(keys, out)}

    @property
    def n_elements(self) -> int:
        return self.mesh.n_elements

    @property
    def n_nodes(self) -> int:
        return self.mesh.n_nodes

    @property
    def n_quad(self) -> int:
        return int(self.element.quad_weights.shape[0])

    def _nodal(self, values):
        """(N,) or (N, ...) -> (device tensor (N, nc) contiguous, trailing shape)."""
        t = nat.to_device(values)
        if t.shape[0] != self.n_nodes:
            raise ValueError(f"nodal array must have {self.n_nodes} rows, got {tuple(t.shape)}")
        trailing = tuple(t.shape[1:])
        nc = int(np.prod(trailing)) if trailing else 1
        return t.reshape(self.n_nodes, nc).contiguous(), trailing, nc

    # --------------------------------------------------------------- ops ----

    def get_integration_weights(self):
        """detJ * w_q per element and quadrature point, (E, Q) (operator.py:172-192)."""
        if self.cache_weights and self._w is not None:
            return self._w
        t = nat.torch()
        out = t.empty((self.n_elements, self.n_quad), dtype=t.float64, device=nat.device())
        nat.check(nat.lib().tatva_weights(self.handle, nat.dptr(out), nat.stream()))
        if self.cache_weights:
            object.__setattr__(self, "_w", out)
        return out

    def eval(self, nodal_values):
        """Values at quadrature points, (E, Q, ...) (operator.py:358-377)."""
        u, trailing, nc = self._nodal(nodal_values)
        t = nat.torch()
        out = t.empty((self.n_elements, self.n_quad, nc), dtype=t.float64, device=u.device)
        nat.check(nat.lib().tatva_eval(self.handle, nc, nat.dptr(u), nat.dptr(out), nat.stream()))
        return out.reshape((self.n_elements, self.n_quad) + trailing)

    def grad(self, nodal_values):
        """Gradients at quadrature points, (E, Q, ..., d) (operator.py:379-397)."""
        u, trailing, nc = self._nodal(nodal_values)
        t = nat.torch()
        d = self.element.dim
        out = t.empty((self.n_elements, self.n_quad, nc, d), dtype=t.float64, device=u.device)
        nat.check(nat.lib().tatva_grad(self.handle, nc, nat.dptr(u), nat.dptr(out), nat.stream()))
        return out.reshape((self.n_elements, self.n_quad) + trailing + (d,))

    def quads(self):
        """Quadrature points in physical coordinates (operator.py:503-511)."""
        return self.eval(self.mesh.coords)

    def integrate_per_element(self, arg):
        """Integral per element (operator.py:321-342): numeric, quad array
        (shape[0] == n_elements) or nodal array."""
        return self._integrate(arg, per_element=True)

    def integrate(self, arg):
        """Integral over the mesh (operator.py:307-319)."""
        return self._integrate(arg, per_element=False)

    def _integrate(self, arg, per_element: bool):
        t = nat.torch()
        if isinstance(arg, (int, float, np.number)):
            vals = t.full((self.n_nodes, 1), float(arg), dtype=t.float64, device=nat.device())
            nodal, trailing = 1, ()
        else:
            a = nat.to_device(arg)
            if a.shape[0] == self.n_elements:
                if a.ndim < 2 or a.shape[1] != self.n_quad:
                    raise ValueError("quadrature arrays must be shaped (n_elements, n_quad, ...)")
                trailing = tuple(a.shape[2:])
                vals, nodal = a, 0
            elif a.shape[0] == self.n_nodes:
                trailing = tuple(a.shape[1:])
                vals, nodal = a, 1
            else:
                raise ValueError("argument is neither a nodal nor a quadrature array")
        nc = int(np.prod(trailing)) if trailing else 1
        vals = vals.contiguous()
        dev = nat.device()
        pe = t.empty((self.n_elements, nc), dtype=t.float64, device=dev) if per_element else None
        tot = None if per_element else t.empty((nc,), dtype=t.float64, device=dev)
        nat.check(nat.lib().tatva_integrate(self.handle, nc, nodal, nat.dptr(vals), nat.dptr(pe),
                                            nat.dptr(tot), nat.stream()))
        if per_element:
            return pe.reshape((self.n_elements,) + trailing)
        return tot.reshape(trailing)

    # ------------------------------------------------------------ mapping --

    def map(self, func, *, element_quantity: Sequence[int] = ()):
        """Map a built-in quadrature-point function over elements and quad points.

        Supported ``func``: ``"gradient"``/``"grad"``, ``"interpolate"``/``"eval"``,
        ``"det_jacobian"``, or a built-in density from ``tatva_amd.energies``
        (mapped over nodal values: psi(grad u)).  Arbitrary callables raise.
        """
        from .energies import Density  # noqa: PLC0415

        if element_quantity:
            if isinstance(func, type) and issubclass(func, Density):
                return self._map_density_with_element_quantities(func, tuple(element_quantity))
            raise NotImplementedError(
                "element_quantity is supported with a built-in density class "
                "(tatva_amd.energies): its parameters are the element quantities")
        if isinstance(func, type) and issubclass(func, Density):
            raise ValueError(f"{func.__name__} as a class needs its parameters as element "
                             "quantities (element_quantity=...); pass an instance otherwise")
        if isinstance(func, Density):
            return lambda u: func(self.grad(u), values=self.eval(u) if func.needs_values else None)
        name = func if isinstance(func, str) else getattr(func, "__name__", None)
        if name in ("gradient", "grad"):
            return self.grad
        if name in ("interpolate", "eval"):
            return self.eval
        if name == "det_jacobian":
            def _det(_coords=None):
                w = self.get_integration_weights()
                return w / nat.to_device(self.element.quad_weights)[None, :]
            return _det
        raise NotImplementedError(
            "Operator.map runs built-in kernels only (gradient, interpolate, det_jacobian, "
            "tatva_amd.energies densities); arbitrary Python callables cannot run on the GPU path")

    def _map_density_with_element_quantities(self, cls, element_quantity):
        """``map(DensityClass, element_quantity=idx)(*values)``: the arguments at
        the indices in ``element_quantity`` are per-element values (one per
        element, passed through as tatva/operator.py:253-258 does), in the order
        of the density's parameters; the one other argument is the nodal field.
        Returns psi at every (element, quadrature point)."""
        def _mapped(*values):
            idx = sorted(set(int(i) for i in element_quantity))
            if any(i < 0 or i >= len(values) for i in idx):
                raise ValueError("element_quantity index out of range")
            nodal = [v for i, v in enumerate(values) if i not in idx]
            if len(nodal) != 1:
                raise ValueError("a density maps one nodal field (the state); got "
                                 f"{len(nodal)} nodal arguments")
            quantities = []
            for i in idx:
                q = nat.to_device(values[i]).reshape(-1)
                if q.shape[0] != self.n_elements:
                    raise ValueError(f"element quantity {i} must have {self.n_elements} values")
                quantities.append(q)
            dens = cls(*quantities)
            u = nodal[0]
            return dens(self.grad(u), values=self.eval(u) if dens.needs_values else None)
        return _mapped

    def map_over_elements(self, func, *, element_quantity: Sequence[int] = ()):
        """Per-element map (operator.py:268-305): built-in ``"integrate"`` only
        (quadrature-point or per-element arrays are integrated per element, so
        element quantities pass through unchanged)."""
        name = func if isinstance(func, str) else getattr(func, "__name__", None)
        if name == "integrate":
            return self.integrate_per_element
        raise NotImplementedError(
            "Operator.map_over_elements runs built-in kernels only; arbitrary Python callables "
            "cannot run on the GPU path")

    # ------------------------------------------------------------- energy --

    def energy(self, density, *, external_force=None):
        """Total-energy functional E(u) of a built-in density (see ``energies``)."""
        from .energies import Energy  # noqa: PLC0415
        return Energy(self, density, external_force=external_force)

    # ------------------------------------------------------ interpolation --

    def interpolate(self, arg, points):
        """Nodal values interpolated at physical points (operator.py:399-468):
        the first element whose vertex polygon holds the point (edges
        included), one Newton step to reference coordinates from the first
        quadrature point, then the element interpolation.  2-D (Tri3) meshes.
        Raises RuntimeError when a point lies in no element."""
        t = nat.torch()
        a = nat.to_device(arg)
        if a.shape[0] != self.n_nodes:
            raise ValueError("interpolate needs nodal values (one row per mesh node)")
        trailing = tuple(a.shape[1:])
        nc = int(np.prod(trailing)) if trailing else 1
        p = nat.to_device(points).reshape(-1, 2).contiguous()
        out = t.empty((p.shape[0], nc), dtype=t.float64, device=a.device)
        outside = ctypes.c_int32(0)
        nat.check(nat.lib().tatva_interpolate(self.handle, p.shape[0], nat.dptr(p), nc,
                                              nat.dptr(a.reshape(self.n_nodes, nc).contiguous()),
                                              nat.dptr(out), ctypes.byref(outside), nat.stream()))
        if outside.value:
            raise RuntimeError("Some points are outside the mesh, revise the points")
        return out.reshape((p.shape[0],) + trailing)

    # --------------------------------------------------------- projection --

    def project(self, field_q, colored_matrix=None, lifter=None, *, tol: float = 1e-13):
        """L2 projection of a quadrature-point field onto the nodal space
        (operator.py:523-554, utils.py:118-262): solve M x = b with the
        consistent mass matrix M_ab = int N_a N_b and b_a = int N_a f.

        ``field_q`` is (n_elements, n_quad, ...).  Without ``colored_matrix``
        (or with a one-dof-per-node pattern) every component is a separate
        right-hand side of the scalar mass matrix; a pattern with ``dim`` dofs
        per node couples the ``dim`` trailing components (block-diagonal mass).
        With a ``lifter`` the system is the reference's (utils.py:225-246): test
        functions at the free dofs, the trial field lifted, so the matrix is the
        Jacobian E_free^T M P of the lifted form (a Periodic slave's column added
        to its master's) and the result is lifted (constraint values imposed).
        The linear solves run to ``tol`` relative residual (the reference uses a
        direct solve; the mass matrix is well conditioned): Jacobi-PCG when the
        matrix is symmetric (no lifter, or Fixed constraints only), else
        Jacobi-BiCGSTAB.
        """
        from . import sparse as sp  # noqa: PLC0415
        from .solvers import bicgstab, pcg  # noqa: PLC0415

        t = nat.torch()
        mapped = lifter is not None and lifter.has_periodic
        f = nat.to_device(field_q)
        if f.ndim < 2 or f.shape[0] != self.n_elements or f.shape[1] != self.n_quad:
            raise ValueError("field must be shaped (n_elements, n_quad, ...)")
        value_shape = tuple(f.shape[2:])
        nvals = int(np.prod(value_shape)) if value_shape else 1
        N = self.n_nodes
        if colored_matrix is None:
            dim = 1
            pat = sp.pattern_from_mesh(self.mesh, 1)
            if lifter is not None:
                pat = lifter.adapt_sparsity(pat)
            colored_matrix = sp.ColoredMatrix.from_csr(pat)
        else:
            n_full = lifter.size if lifter is not None else colored_matrix.shape[0]
            if lifter is not None and colored_matrix.shape[0] != lifter.size_reduced:
                raise ValueError("Colored matrix size does not match lifter reduced size. "
                                 f"Expected {lifter.size_reduced}, got {colored_matrix.shape[0]}")
            if n_full % N:
                raise ValueError("the pattern size is not a multiple of the number of nodes")
            dim = n_full // N
        if dim > 1 and (nvals % dim or (value_shape and value_shape[-1] != dim)):
            raise ValueError(f"a {dim}-dof-per-node pattern needs fields with trailing size {dim}")
        cm = colored_matrix
        L = nat.lib()
        ws = t.empty(int(L.tatva_colored_hessian_workspace(self.handle, dim)), dtype=t.uint8,
                     device=f.device)
        data = t.empty(cm.nnz, dtype=t.float64, device=f.device)
        if mapped:
            # rows: each free test dof alone; columns: through the lift's read map
            n_red = lifter.size_reduced
            row_ptr = t.arange(n_red + 1, dtype=t.int32, device=f.device)
            row_list = lifter.free_device()
            flag = _periodic_serial_flag(self.mesh.elements, lifter, dim)
            nat.check(L.tatva_mass_matrix_mapped(
                self.handle, dim, nat.dptr(lifter.read_map_device()), nat.dptr(row_ptr),
                nat.dptr(row_list), cm.shape[0], nat.dptr(cm.colors), cm.n_colors,
                nat.dptr(cm.indptr), nat.dptr(cm.indices), cm.nnz, flag, nat.dptr(data),
                nat.dptr(ws), ws.numel(), nat.stream()))
        else:
            dmap = lifter.dof_map_device() if lifter is not None else None
            r2f = lifter.free_device() if lifter is not None else None
            nat.check(L.tatva_mass_matrix(self.handle, dim, nat.dptr(dmap), nat.dptr(r2f),
                                          cm.shape[0], nat.dptr(cm.colors), cm.n_colors,
                                          nat.dptr(cm.indptr), nat.dptr(cm.indices), cm.nnz,
                                          nat.dptr(data), nat.dptr(ws), ws.numel(), nat.stream()))
        from dataclasses import replace  # noqa: PLC0415
        M = replace(cm, data=data)
        vals = f.reshape(self.n_elements, self.n_quad, nvals).contiguous()
        b = t.empty((N, nvals), dtype=t.float64, device=f.device)
        nat.check(L.tatva_project_rhs(self.handle, nvals, nat.dptr(vals), nat.dptr(b),
                                      nat.stream()))
        # right-hand sides in solver layout: (n_rhs, N * dim)
        rhs = b.reshape(N, nvals // dim, dim).permute(1, 0, 2).reshape(nvals // dim, N * dim)
        out = t.empty_like(rhs)
        for k in range(rhs.shape[0]):
            bk = rhs[k].contiguous()
            if lifter is not None:
                bk = lifter.reduce(bk)      # test functions injected at the free dofs
            if mapped:
                xk, info = bicgstab(M, bk, tol=tol)
            else:
                xk, info = pcg(M, bk, tol=tol, block_size=dim if dim in (2, 3) else None)
            if lifter is not None:
                xk = lifter.lift_from_zeros(xk)
            out[k] = xk
        res = out.reshape(nvals // dim, N, dim).permute(1, 0, 2).reshape((N,) + value_shape)
        return res.contiguous()

    def close(self) -> None:
        """Release the device-resident mesh now (otherwise freed with the Operator)."""
        if self._dev is not None:
            self._dev.close()
            object.__setattr__(self, "_dev", None)


def _periodic_serial_flag(elements, lifter, nc) -> int:
    """1 when some element holds two dofs that read the same reduced index (a
    periodic slave and its master in one element): the mapped decompression
    then accumulates an element's columns one at a time."""
    el = np.asarray(elements, dtype=np.int64)
    dofs = (el[:, :, None] * nc + np.arange(nc)).reshape(el.shape[0], -1)
    rm = lifter.read_map[dofs]
    rs = np.sort(np.where(rm >= 0, rm, -1 - np.arange(rm.shape[1])[None, :]), axis=1)
    return int(np.any(rs[:, 1:] == rs[:, :-1]))


__all__ = ["Operator"]
