"""Built-in energy densities and the energy functional E(u) with its derivatives.

In tatva a user writes the energy as JAX code, e.g. (tests/test_sparse.py:50-55)::

    def total_energy(u_flat):
        return op.integrate(strain_energy(op.grad(u_flat.reshape(-1, 2)), mu, lmbda))

and obtains the residual with ``jax.grad`` and Hessian-vector products with
``jax.jvp(jax.grad(E), (u,), (v,))`` (README.md:93).  A HIP kernel cannot
differentiate a Python callable, so the densities are built in, with
hand-derived first and second derivatives in the kernels:

    E = op.energy(LinearElasticity(mu, lmbda))    # E(u) -> scalar
    r = E.grad(u)                                  # jax.grad(E)(u)
    y = E.hvp(u, v)                                # jax.jvp(jax.grad(E), (u,), (v,))[1]
    psi = LinearElasticity(mu, lmbda)(op.grad(u))  # density at quad points

``E.reduced(lifter)`` is the energy of the free dofs, E(lifter.lift_from_zeros(u)),
with residual ``lifter.reduce_adjoint(grad E)`` (the pattern of
tests/test_sparse_tracer.py:140-149).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat


def _param(x):
    """A density parameter: a float, or per-element values (a 1-D array)."""
    if np.ndim(x) == 0:
        return float(x)
    a = np.ascontiguousarray(nat_host(x), dtype=np.float64)
    if a.ndim != 1:
        raise ValueError("per-element parameters must be 1-D arrays (one value per element)")
    return a


def nat_host(a):
    t = nat.torch()
    if isinstance(a, t.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


class Density:
    """Base class of the built-in densities (FP64 parameters).

    Each parameter is a float, or a 1-D array with one value per element: the
    element quantities that tatva passes through ``Operator.map(...,
    element_quantity=...)`` (tatva/operator.py:229,256), e.g. two materials in
    one mesh.  Per-element parameters travel to the kernels as one (E, np)
    device array."""

    kind: int = 0
    needs_values = False

    def params(self):
        raise NotImplementedError

    def ncomp(self, dim: int) -> int:
        return dim

    @property
    def per_element(self) -> bool:
        return any(np.ndim(p) for p in self.params())

    def n_elements(self):
        """Number of elements of per-element parameters (None when uniform)."""
        sizes = {np.shape(p)[0] for p in self.params() if np.ndim(p)}
        if len(sizes) > 1:
            raise ValueError("per-element parameters of different lengths")
        return sizes.pop() if sizes else None

    def element_params(self):
        """(E, np) FP64 device array of the per-element parameters (cached)."""
        dev = getattr(self, "_elem_dev", None)
        if dev is None:
            E = self.n_elements()
            cols = [np.broadcast_to(np.asarray(p, dtype=np.float64), (E,)) for p in self.params()]
            dev = nat.to_device(np.ascontiguousarray(np.stack(cols, axis=1)))
            self._elem_dev = dev
        return dev

    def descriptor(self, points_per_element: int = 0) -> nat.Density:
        if self.per_element:
            ep = self.element_params()
            return nat.Density(self.kind, int(points_per_element), 0.0, 0.0, 0.0, nat.dptr(ep))
        p = list(self.params()) + [0.0] * 3
        return nat.Density(self.kind, 0, float(p[0]), float(p[1]), float(p[2]), None)

    def __call__(self, grad_u, values=None):
        """psi at quadrature points from gradients (..., ncomp, d) -> (...)."""
        g = nat.to_device(grad_u)
        d = int(g.shape[-1])
        vals = None
        if self.needs_values:
            # scalar field: values (...) at the same points as grad (..., d)
            if values is None:
                raise ValueError(f"{type(self).__name__} needs the field values at quad points")
            vals = nat.to_device(values)
            lead = tuple(vals.shape)
            npts = int(np.prod(lead)) if lead else 1
            if g.numel() != npts * d:
                raise ValueError("values and gradients do not describe the same points")
            vals = vals.reshape(npts).contiguous()
        else:
            if tuple(g.shape[-2:]) != (d, d):
                raise ValueError(f"gradient array must end in ({d}, {d}), got {tuple(g.shape)}")
            lead = tuple(g.shape[:-2])
            npts = int(np.prod(lead)) if lead else 1
        t = nat.torch()
        out = t.empty(lead, dtype=t.float64, device=g.device)
        ppe = 0
        if self.per_element:
            # points (E, Q, ...) in element order
            if not lead or lead[0] != self.n_elements():
                raise ValueError(f"per-element parameters of {self.n_elements()} elements need "
                                 "points shaped (n_elements, ...)")
            ppe = npts // lead[0]
        desc = self.descriptor(ppe)
        nat.check(nat.lib().tatva_density_eval(d, npts, ctypes.byref(desc), nat.dptr(g.contiguous()),
                                               nat.dptr(vals), nat.dptr(out), nat.stream()))
        return out


class Poisson(Density):
    """psi = 1/2 |grad u|^2 - f u  (scalar field)."""

    kind = nat.POISSON
    needs_values = True

    def __init__(self, f=1.0):
        self.f = _param(f)

    def params(self):
        return (self.f,)

    def ncomp(self, dim: int) -> int:
        return 1

    def __repr__(self):
        return f"Poisson(f={self.f})"


class LinearElasticity(Density):
    """psi = 1/2 sigma:eps, sigma = 2 mu eps + lambda tr(eps) I (tests/test_sparse.py:20-38)."""

    kind = nat.LINEAR_ELASTIC

    def __init__(self, mu, lmbda):
        self.mu, self.lmbda = _param(mu), _param(lmbda)

    @classmethod
    def from_youngs_poisson(cls, E: float, nu: float) -> "LinearElasticity":
        """mu = E/(2(1+nu)), lambda = E nu/((1+nu)(1-2nu))
        (tests/test_colored_matrix_benchmark.py:36-45)."""
        return cls(E / 2 / (1 + nu), E * nu / ((1 + nu) * (1 - 2 * nu)))

    def params(self):
        return (self.mu, self.lmbda)

    def __repr__(self):
        return f"LinearElasticity(mu={self.mu}, lmbda={self.lmbda})"


class NeoHookean(Density):
    """psi = mu/2 (tr C - d - 2 ln J) + lambda/2 (ln J)^2, F = I + grad u
    (tests/test_sparse_tracer.py:103-115)."""

    kind = nat.NEO_HOOKEAN

    def __init__(self, mu, lmbda):
        self.mu, self.lmbda = _param(mu), _param(lmbda)

    def params(self):
        return (self.mu, self.lmbda)

    def __repr__(self):
        return f"NeoHookean(mu={self.mu}, lmbda={self.lmbda})"


# --------------------------------------------------------------------------


class Residual:
    """The gradient dE/du of an energy, as a callable (``jax.grad(E)``)."""

    def __init__(self, energy):
        self.energy = energy

    def __call__(self, u):
        return self.energy.residual(u)

    def jvp(self, u, v):
        """(r(u), H(u) v), as ``jax.jvp(jax.grad(E), (u,), (v,))``."""
        return self.energy.residual(u), self.energy.hvp(u, v)


class Energy:
    """E(u) = integral of psi(grad u) over the mesh - f_ext . u."""

    def __init__(self, op, density: Density, *, external_force=None):
        self.op = op
        self.density = density
        self.ncomp = density.ncomp(op.element.dim)
        self.size = op.n_nodes * self.ncomp
        E = density.n_elements()
        if E is not None and E != op.n_elements:
            raise ValueError(f"per-element parameters of {E} elements for a mesh of "
                             f"{op.n_elements} elements")
        self._desc = density.descriptor()
        self.f_ext = None if external_force is None else nat.to_device(external_force).reshape(-1)
        if self.f_ext is not None and self.f_ext.shape[0] != self.size:
            raise ValueError(f"external_force must have {self.size} entries")

    def _vec(self, u, dtype=None):
        t = nat.torch()
        x = nat.to_device(u, dtype or t.float64).reshape(-1)
        if x.shape[0] != self.size:
            raise ValueError(f"state vector must have {self.size} entries, got {x.shape[0]}")
        return x.contiguous()

    @property
    def grad(self) -> Residual:
        return Residual(self)

    def __call__(self, u):
        t = nat.torch()
        x = self._vec(u)
        out = t.empty((), dtype=t.float64, device=x.device)
        nat.check(nat.lib().tatva_energy(self.op.handle, ctypes.byref(self._desc), nat.dptr(x),
                                         nat.dptr(self.f_ext), nat.dptr(out), nat.stream()))
        return out

    def residual(self, u, *, dof_map=None, n_out=None):
        t = nat.torch()
        x = self._vec(u)
        r = t.empty((n_out or self.size,), dtype=t.float64, device=x.device)
        nat.check(nat.lib().tatva_residual(self.op.handle, ctypes.byref(self._desc), nat.dptr(x),
                                           nat.dptr(self.f_ext), nat.dptr(dof_map), nat.dptr(r),
                                           nat.stream()))
        return r

    def hvp(self, u, v, *, dof_map=None, n_out=None):
        """H(u) v in FP64 (or FP32 when both u and v are float32 tensors)."""
        t = nat.torch()
        f32 = isinstance(u, t.Tensor) and isinstance(v, t.Tensor) and \
            u.dtype == t.float32 and v.dtype == t.float32
        dt = t.float32 if f32 else t.float64
        x = self._vec(u, dt)
        n = n_out or self.size
        y = nat.to_device(v, dt).reshape(-1).contiguous()
        if y.shape[0] != n:
            raise ValueError(f"tangent must have {n} entries")
        out = t.empty((n,), dtype=dt, device=x.device)
        fn = nat.lib().tatva_hvp_f32 if f32 else nat.lib().tatva_hvp
        nat.check(fn(self.op.handle, ctypes.byref(self._desc), nat.dptr(x), nat.dptr(y),
                     nat.dptr(dof_map), nat.dptr(out), nat.stream()))
        return out

    def hessian_diag(self, u, *, dof_map=None, n_out=None, dtype=None):
        t = nat.torch()
        dt = dtype or t.float64
        x = self._vec(u, dt)
        out = t.empty((n_out or self.size,), dtype=dt, device=x.device)
        fn = nat.lib().tatva_hessian_diag_f32 if dt == t.float32 else nat.lib().tatva_hessian_diag
        nat.check(fn(self.op.handle, ctypes.byref(self._desc), nat.dptr(x), nat.dptr(dof_map),
                     nat.dptr(out), nat.stream()))
        return out

    def linearize(self, u):
        """(r(u), v -> H(u) v), as ``jax.linearize(jax.grad(E), u)``."""
        x = self._vec(u)
        return self.residual(x), (lambda v: self.hvp(x, v))

    def reduced(self, lifter) -> "ReducedEnergy":
        return ReducedEnergy(self, lifter)


class ReducedEnergy:
    """E_free(u) = E(lifter.lift_from_zeros(u)) on the free dofs of a Lifter."""

    def __init__(self, energy: Energy, lifter):
        if lifter.size != energy.size:
            raise ValueError(f"lifter size {lifter.size} != energy size {energy.size}")
        lifter._require_device_kernels()
        self.energy = energy
        self.lifter = lifter
        self.size = lifter.size_reduced
        self.op = energy.op
        self.density = energy.density
        self.ncomp = energy.ncomp

    @property
    def grad(self) -> Residual:
        return Residual(self)

    def lift(self, u_red):
        return self.lifter.lift_from_zeros(u_red)

    def __call__(self, u_red):
        return self.energy(self.lift(u_red))

    def residual(self, u_red):
        return self.residual_full_state(self.lift(u_red))

    def residual_full_state(self, u_full):
        """Reduced residual at an already lifted state: the kernels write the free
        rows directly; with Periodic constraints the full residual goes through
        reduce_adjoint (a slave's row is added to its master's)."""
        lf = self.lifter
        if lf.has_periodic:
            return lf.reduce_adjoint(self.energy.residual(u_full))
        return self.energy.residual(u_full, dof_map=lf.dof_map_device(), n_out=self.size)

    def hvp(self, u_red, v_red):
        return self.hvp_full_state(self.lift(u_red), v_red)

    def hvp_full_state(self, u_full, v_red):
        lf = self.lifter
        if lf.has_periodic:
            t = nat.torch()
            if isinstance(v_red, t.Tensor) and v_red.dtype == t.float32:
                raise NotImplementedError("FP32 HVP with Periodic constraints is not supported")
            y = self.energy.hvp(u_full, lf.lift_linear(v_red))
            return lf.reduce_adjoint(y)
        return self.energy.hvp(u_full, v_red, dof_map=lf.dof_map_device(), n_out=self.size)

    def hessian_diag(self, u_red, dtype=None):
        """Jacobi diagonal.  With Periodic constraints it sums the full diagonal
        entries of the dofs reading each reduced index (exact unless a slave and
        its master share an element)."""
        lf = self.lifter
        uf = self.lift(u_red)
        if lf.has_periodic:
            d = lf.reduce_adjoint(self.energy.hessian_diag(uf))
            return d if dtype is None else d.to(dtype)
        return self.energy.hessian_diag(uf, dof_map=lf.dof_map_device(), n_out=self.size,
                                        dtype=dtype)

    def linearize(self, u_red):
        uf = self.lift(u_red)
        return self.residual_full_state(uf), (lambda v: self.hvp_full_state(uf, v))


# ------------------------------------------------- jax-style entry points --


def grad(energy):
    """``jax.grad(E)`` for a built-in energy."""
    if not isinstance(energy, (Energy, ReducedEnergy)):
        raise NotImplementedError("grad() differentiates built-in energies only")
    return energy.grad


def jvp(fn, primals, tangents):
    """``jax.jvp(jax.grad(E), (u,), (v,))`` -> (r(u), H(u) v)."""
    if not isinstance(fn, Residual):
        raise NotImplementedError("jvp() needs the residual of a built-in energy (E.grad)")
    (u,), (v,) = primals, tangents
    return fn.jvp(u, v)


def linearize(fn, u):
    """``jax.linearize(jax.grad(E), u)`` -> (r(u), v -> H(u) v)."""
    if not isinstance(fn, Residual):
        raise NotImplementedError("linearize() needs the residual of a built-in energy (E.grad)")
    return fn.energy.linearize(u)


__all__ = ["Density", "Poisson", "LinearElasticity", "NeoHookean", "Energy", "ReducedEnergy",
           "Residual", "grad", "jvp", "linearize"]
