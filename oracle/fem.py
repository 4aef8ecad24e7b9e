"""NumPy/SciPy FP64 restatement of tatva's hot path.  TEST INFRASTRUCTURE ONLY.

Every function cites the reference file:line it restates (paths relative to the
tatva repository).  Arrays follow the reference layouts: coords (N, d) f64,
elements (E, nne) int32, nodal dofs interleaved node-major (dof = node*ncomp+c),
gradients ``grad[..., i, j] = d u_i / d x_j``.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sps

# --------------------------------------------------------------------------
# Meshes
# --------------------------------------------------------------------------


def unit_square_tri3(nx: int, ny: int):
    """``Mesh.unit_square`` -> ``_rectangle_triangular`` (mesh.py:146-156,180-205).

    Node id ``i*(ny+1)+j`` from ``meshgrid(x, y, indexing='ij')``; each cell
    (i, j) is split into ``[n0, n1, n3]`` and ``[n0, n3, n2]``.
    """
    x = np.linspace(0.0, 1.0, nx + 1)
    y = np.linspace(0.0, 1.0, ny + 1)
    xv, yv = np.meshgrid(x, y, indexing="ij")
    coords = np.stack([xv.ravel(), yv.ravel()], axis=-1)
    elems = []
    for i in range(nx):
        for j in range(ny):
            n0 = i * (ny + 1) + j
            n1 = (i + 1) * (ny + 1) + j
            n2 = i * (ny + 1) + j + 1
            n3 = (i + 1) * (ny + 1) + j + 1
            elems.append([n0, n1, n3])
            elems.append([n0, n3, n2])
    return coords, np.asarray(elems, dtype=np.int32)


def cube_hex8(n: int, length: float = 1.0):
    """Structured Hex8 cube [0, L]^3 with n cells per side.

    Node id ``i + (n+1) j + (n+1)^2 k``; element e = i + n j + n^2 k with nodes in
    the ``Hexahedron8`` reference order (element/base.py:478-491):
    (-,-,-), (+,-,-), (+,+,-), (-,+,-), (-,-,+), (+,-,+), (+,+,+), (-,+,+).
    """
    m = n + 1
    t = np.linspace(0.0, length, m)
    zz, yy, xx = np.meshgrid(t, t, t, indexing="ij")
    coords = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=-1)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    n0 = (i + m * j + m * m * k).ravel()
    sx, sy, sz = 1, m, m * m
    conn = np.stack(
        [n0, n0 + sx, n0 + sx + sy, n0 + sy,
         n0 + sz, n0 + sz + sx, n0 + sz + sx + sy, n0 + sz + sy], axis=-1)
    return coords, conn.astype(np.int32)


def box_tet4(nx: int, ny: int, nz: int, lengths=(1.0, 1.0, 1.0), jitter=0.0, seed=0):
    """6-tets-per-cell box following ``tests/test_sparse_tracer.py:29-70``.

    Node numbering ``i + (nx+1) j + (nx+1)(ny+1) k``; cells ordered k, j, i
    (i fastest); per cell the six tets t1..t6 of the reference (all share the
    n0-n7 diagonal).  The box here is [0, Lx] x [0, Ly] x [0, Lz].  With
    ``jitter > 0`` interior nodes are moved by U(-jitter*h, jitter*h) per
    coordinate drawn from ``numpy.random.default_rng(seed)``.
    """
    x = np.linspace(0.0, lengths[0], nx + 1)
    y = np.linspace(0.0, lengths[1], ny + 1)
    z = np.linspace(0.0, lengths[2], nz + 1)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    coords = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    sx, sy, sz = 1, nx + 1, (nx + 1) * (ny + 1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (i * sx + j * sy + k * sz).ravel()
    n1 = n0 + sx
    n2 = n0 + sy
    n3 = n2 + sx
    n4 = n0 + sz
    n5 = n4 + sx
    n6 = n4 + sy
    n7 = n6 + sx
    tets = np.stack([
        np.stack([n0, n1, n3, n7], -1), np.stack([n0, n1, n7, n5], -1),
        np.stack([n0, n5, n7, n4], -1), np.stack([n0, n3, n2, n7], -1),
        np.stack([n0, n2, n6, n7], -1), np.stack([n0, n6, n4, n7], -1)], axis=1)
    elems = tets.reshape(-1, 4).astype(np.int32)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        h = np.array([lengths[0] / nx, lengths[1] / ny, lengths[2] / nz])
        delta = rng.uniform(-jitter, jitter, size=coords.shape) * h
        lo = np.isclose(coords, 0.0)
        hi = np.isclose(coords, np.asarray(lengths))
        interior = ~(lo | hi).any(axis=1)
        coords = coords + delta * interior[:, None]
    return coords, elems


# --------------------------------------------------------------------------
# Elements (element/base.py)
# --------------------------------------------------------------------------


class Element:
    dim: int
    nne: int

    def __init__(self, quad_points=None, quad_weights=None):
        """``Element.__init__`` (element/base.py:37-51)."""
        if quad_points is not None and quad_weights is not None:
            self.quad_points = np.asarray(quad_points, dtype=np.float64)
            self.quad_weights = np.asarray(quad_weights, dtype=np.float64)
        else:
            self.quad_points, self.quad_weights = self.default_quadrature()


class Tri3(Element):
    """element/base.py:245-265."""

    dim, nne = 2, 3
    ref_nodes = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])

    def default_quadrature(self):
        return np.array([[1.0 / 3, 1.0 / 3]]), np.array([0.5])

    def N(self, xi):
        return np.array([1.0 - xi[0] - xi[1], xi[0], xi[1]])

    def dN(self, xi):
        return np.array([[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]]).T


class Tet4(Element):
    """``Tetrahedron4`` (element/base.py:448-472)."""

    dim, nne = 3, 4
    ref_nodes = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1, 0], [0.0, 0, 1]])

    def default_quadrature(self):
        return np.array([[0.25, 0.25, 0.25]]), np.array([1.0 / 6])

    def N(self, xi):
        return np.array([1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]])

    def dN(self, xi):
        return np.array([[-1.0, -1, -1], [1.0, 0, 0], [0.0, 1, 0], [0.0, 0, 1]]).T


class Hex8(Element):
    """``Hexahedron8`` (element/base.py:475-568)."""

    dim, nne = 3, 8
    ref_nodes = np.array([[-1.0, -1, -1], [1.0, -1, -1], [1.0, 1, -1], [-1.0, 1, -1],
                          [-1.0, -1, 1], [1.0, -1, 1], [1.0, 1, 1], [-1.0, 1, 1]])

    def default_quadrature(self):
        a = 1.0 / np.sqrt(3.0)
        s = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                      [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
        return a * s, np.ones(8)

    def N(self, xi):
        r = self.ref_nodes
        return 0.125 * (1 + r[:, 0] * xi[0]) * (1 + r[:, 1] * xi[1]) * (1 + r[:, 2] * xi[2])

    def dN(self, xi):
        r = self.ref_nodes
        fx = 1 + r[:, 0] * xi[0]
        fy = 1 + r[:, 1] * xi[1]
        fz = 1 + r[:, 2] * xi[2]
        return 0.125 * np.stack([r[:, 0] * fy * fz, fx * r[:, 1] * fz, fx * fy * r[:, 2]])


ELEMENTS = {"tri3": Tri3, "tet4": Tet4, "hex8": Hex8}


# --------------------------------------------------------------------------
# Operator (operator.py)
# --------------------------------------------------------------------------


def jacobians(elem: Element, coords, elements):
    """``Element.get_jacobian`` over all elements and quadrature points.

    ``J = dNdr @ X`` and signed ``det(J)`` (element/base.py:90-93); returns
    J (E, Q, d, d) and detJ (E, Q).
    """
    X = coords[elements]                                  # (E, nne, d)
    dN = np.stack([elem.dN(xi) for xi in elem.quad_points])   # (Q, d, nne)
    J = np.einsum("qdn,enk->eqdk", dN, X)
    return J, np.linalg.det(J)


def integration_weights(elem, coords, elements):
    """``Operator.get_integration_weights`` (operator.py:172-192): detJ * w_q."""
    _, detJ = jacobians(elem, coords, elements)
    return detJ * elem.quad_weights[None, :]


def shape_gradients(elem, coords, elements):
    """dNdX = inv(J) @ dNdr per (e, q) (element/base.py:99-115): (E, Q, d, nne)."""
    J, _ = jacobians(elem, coords, elements)
    dN = np.stack([elem.dN(xi) for xi in elem.quad_points])
    return np.einsum("eqij,qjn->eqin", np.linalg.inv(J), dN)


def op_eval(elem, coords, elements, nodal):
    """``Operator.eval`` (operator.py:358-377): (E, Q, ...) values at quad points."""
    N = np.stack([elem.N(xi) for xi in elem.quad_points])   # (Q, nne)
    return np.einsum("qn,en...->eq...", N, nodal[elements])


def op_grad(elem, coords, elements, nodal):
    """``Operator.grad`` (operator.py:379-397): (E, Q, ..., d)."""
    B = shape_gradients(elem, coords, elements)
    return np.einsum("eqdn,en...->eq...d", B, nodal[elements])


def integrate_per_element(elem, coords, elements, arg):
    """``Operator.integrate_per_element`` (operator.py:321-356).

    Numeric -> eval of a constant; ``shape[0] == E`` -> quad field; otherwise a
    nodal field evaluated at the quadrature points first.
    """
    w = integration_weights(elem, coords, elements)
    E = elements.shape[0]
    if np.isscalar(arg):
        vals = op_eval(elem, coords, elements, np.full(coords.shape[0], float(arg)))
    else:
        arg = np.asarray(arg, dtype=np.float64)
        vals = arg if arg.shape[0] == E else op_eval(elem, coords, elements, arg)
    return np.einsum("eq...,eq->e...", vals, w)


def integrate(elem, coords, elements, arg):
    """``Operator.integrate`` (operator.py:307-319)."""
    return integrate_per_element(elem, coords, elements, arg).sum(axis=0)


# --------------------------------------------------------------------------
# Energy densities (tests/test_sparse.py:20-38, tests/test_sparse_tracer.py:103-115)
# --------------------------------------------------------------------------


def _inv_T(F):
    return np.swapaxes(np.linalg.inv(F), -1, -2)


def _pe(p, like):
    """A density parameter against an (E, ...) array: a scalar as is, per-element
    values (E,) with trailing unit axes (the element quantities of
    Operator.map(..., element_quantity=...), tatva/operator.py:229,256)."""
    if np.ndim(p) == 0:
        return p
    return np.asarray(p, dtype=np.float64).reshape((-1,) + (1,) * (np.ndim(like) - 1))


@dataclass
class Poisson:
    """psi = 1/2 |grad u|^2 - f u (scalar field, ncomp = 1)."""

    f: float = 1.0
    scalar = True

    def psi(self, G, uq):
        return 0.5 * np.sum(G * G, axis=(-2, -1)) - _pe(self.f, uq[..., 0]) * uq[..., 0]

    def P(self, G):
        return G

    def dP(self, G, dG):
        return dG


@dataclass
class LinearElasticity:
    """psi = 1/2 sigma:eps, sigma = 2 mu eps + lambda tr(eps) I (tests/test_sparse.py:20-38)."""

    mu: float
    lmbda: float
    scalar = False

    def _sig(self, G):
        d = G.shape[-1]
        eps = 0.5 * (G + np.swapaxes(G, -1, -2))
        tr = np.trace(eps, axis1=-2, axis2=-1)
        return (2 * _pe(self.mu, G) * eps + _pe(self.lmbda, G) * tr[..., None, None] * np.eye(d),
                eps)

    def psi(self, G, uq=None):
        sig, eps = self._sig(G)
        return 0.5 * np.sum(sig * eps, axis=(-2, -1))

    def P(self, G):
        return self._sig(G)[0]

    def dP(self, G, dG):
        return self._sig(dG)[0]


@dataclass
class NeoHookean:
    """psi = mu/2 (tr C - d - 2 ln J) + lambda/2 (ln J)^2 with F = I + grad u
    (tests/test_sparse_tracer.py:103-115; d = 3 there)."""

    mu: float
    lmbda: float
    scalar = False

    def psi(self, G, uq=None):
        d = G.shape[-1]
        F = np.eye(d) + G
        J = np.linalg.det(F)
        lnJ = np.log(J)
        I1 = np.sum(F * F, axis=(-2, -1))
        return 0.5 * _pe(self.mu, lnJ) * (I1 - d - 2 * lnJ) + 0.5 * _pe(self.lmbda, lnJ) * lnJ ** 2

    def P(self, G):
        d = G.shape[-1]
        F = np.eye(d) + G
        lnJ = np.log(np.linalg.det(F))
        FiT = _inv_T(F)
        return _pe(self.mu, F) * (F - FiT) + _pe(self.lmbda, F) * lnJ[..., None, None] * FiT

    def dP(self, G, dG):
        """dP = mu dF + (mu - lambda lnJ) F^-T dF^T F^-T + lambda tr(F^-1 dF) F^-T."""
        d = G.shape[-1]
        F = np.eye(d) + G
        lnJ = np.log(np.linalg.det(F))
        Fi = np.linalg.inv(F)
        FiT = np.swapaxes(Fi, -1, -2)
        tr = np.einsum("...ij,...ji->...", Fi, dG)
        t2 = FiT @ np.swapaxes(dG, -1, -2) @ FiT
        c1 = _pe(self.mu, lnJ) - _pe(self.lmbda, lnJ) * lnJ
        return (_pe(self.mu, dG) * dG + c1[..., None, None] * t2
                + _pe(self.lmbda, dG) * tr[..., None, None] * FiT)


# --------------------------------------------------------------------------
# Energy, residual (= jax.grad), HVP (= jax.jvp(jax.grad)) and element stiffness
# --------------------------------------------------------------------------


class Problem:
    """A mesh + element + density; vectors are flat node-major (N*ncomp,)."""

    def __init__(self, elem, coords, elements, density, f_ext=None):
        self.elem, self.coords, self.elements, self.density = elem, coords, elements, density
        self.nc = 1 if density.scalar else elem.dim
        self.n = coords.shape[0] * self.nc
        self.w = integration_weights(elem, coords, elements)                # (E, Q)
        self.B = shape_gradients(elem, coords, elements)                   # (E, Q, d, nne)
        self.Nq = np.stack([elem.N(xi) for xi in elem.quad_points])        # (Q, nne)
        self.f_ext = None if f_ext is None else np.asarray(f_ext, dtype=np.float64)

    def _ue(self, u):
        return u.reshape(-1, self.nc)[self.elements]                      # (E, nne, nc)

    def grad_q(self, u):
        return np.einsum("eqdn,enc->eqcd", self.B, self._ue(u))            # (E, Q, nc, d)

    def energy(self, u):
        """``Operator.integrate(psi(Operator.grad(u)))`` (+ Poisson source via eval)."""
        G = self.grad_q(u)
        uq = np.einsum("qn,enc->eqc", self.Nq, self._ue(u))
        e = np.sum(self.density.psi(G, uq) * self.w)
        if self.f_ext is not None:
            e -= float(self.f_ext @ u)
        return e

    def _assemble(self, re):
        out = np.zeros((self.coords.shape[0], self.nc))
        np.add.at(out, self.elements, re)
        return out.ravel()

    def residual(self, u):
        """r = dE/du: scatter-add transpose of the gather (operator.py:221)."""
        P = self.density.P(self.grad_q(u))                                  # (E,Q,nc,d)
        re = np.einsum("eq,eqcd,eqdn->enc", self.w, P, self.B)
        if isinstance(self.density, Poisson):
            src = np.einsum("eq,qn->en", self.w, self.Nq)[..., None]
            re -= _pe(self.density.f, src) * src
        r = self._assemble(re)
        if self.f_ext is not None:
            r -= self.f_ext
        return r

    def hvp(self, u, v):
        """y = H(u) v (``jax.jvp(jax.grad(E), (u,), (v,))[1]``, README.md:93)."""
        G = self.grad_q(u)
        dG = self.grad_q(v)
        dP = self.density.dP(G, dG)
        ye = np.einsum("eq,eqcd,eqdn->enc", self.w, dP, self.B)
        return self._assemble(ye)

    def element_stiffness(self, u):
        """K_e[(a,i),(b,k)] = sum_q w B_ja dP_ij(e_(b,k)) (E, nne*nc, nne*nc)."""
        E, Q, d, nne = self.B.shape
        nc = self.nc
        G = self.grad_q(u)
        K = np.zeros((E, nne * nc, nne * nc))
        for b in range(nne):
            for k in range(nc):
                dG = np.zeros((E, Q, nc, d))
                dG[:, :, k, :] = self.B[:, :, :, b]
                dP = self.density.dP(G, dG)
                col = np.einsum("eq,eqcd,eqdn->enc", self.w, dP, self.B)
                K[:, :, b * nc + k] = col.reshape(E, nne * nc)
        return K

    def hessian_diag(self, u):
        K = self.element_stiffness(u)
        de = np.diagonal(K, axis1=1, axis2=2).reshape(-1, self.elements.shape[1], self.nc)
        return self._assemble(de)


# --------------------------------------------------------------------------
# Sparsity pattern (sparse/_extraction.py:37-102)
# --------------------------------------------------------------------------


def pattern_from_mesh(elements, n_nodes, ndpn):
    """``_create_sparse_structure`` restated: all element dof pairs, deduplicated and
    sorted through the linearised index ``row*ncols+col`` (sparse/_extraction.py:37-87)."""
    el = np.asarray(elements, dtype=np.int64)
    E, nne = el.shape
    nd = nne * ndpn
    dofs = (el[..., None] * ndpn + np.arange(ndpn, dtype=np.int64)).reshape(E, nd)
    rows = np.repeat(dofs, nd, axis=1).ravel()
    cols = np.tile(dofs, (1, nd)).ravel()
    n = n_nodes * ndpn
    lin = np.unique(rows * n + cols)
    r, c = lin // n, lin % n
    return sps.csr_matrix((np.ones(r.shape[0], dtype=np.int8), (r, c)), shape=(n, n))


# --------------------------------------------------------------------------
# Lifter / Fixed / Periodic (lifter/base.py, lifter/constraints.py)
# --------------------------------------------------------------------------


class Fixed:
    def __init__(self, dofs, values=0.0):
        self.dofs = np.asarray(dofs, dtype=np.int64).ravel()
        self.values = values

    def lift(self, u):            # constraints.py:312-314
        u[self.dofs] = self.values
        return u

    def transpose(self, r):       # constraints.py:316-318
        r[self.dofs] = 0.0
        return r

    def augment(self, S):
        return S


class Periodic:
    def __init__(self, dofs, master_dofs):
        self.dofs = np.asarray(dofs, dtype=np.int64).ravel()
        self.master = np.asarray(master_dofs, dtype=np.int64).ravel()

    def lift(self, u):            # constraints.py:214-216
        u[self.dofs] = u[self.master]
        return u

    def transpose(self, r):       # constraints.py:218-220
        np.add.at(r, self.master, r[self.dofs])
        r[self.dofs] = 0.0
        return r

    def augment(self, S):         # constraints.py:194-212: M^T S M
        n = S.shape[0]
        rows = np.concatenate([np.arange(n), self.dofs])
        cols = np.concatenate([np.arange(n), self.master])
        M = sps.csr_matrix((np.ones(rows.size, dtype=np.int8), (rows, cols)), shape=(n, n))
        out = (M.T @ (S.astype(np.int8) @ M)).tocsr()
        out.data = np.ones_like(out.data, dtype=np.int8)
        return out


class Lifter:
    """lifter/base.py:55-331 for Fixed and Periodic constraints."""

    def __init__(self, size, *constraints):
        self.size = size
        self.constraints = constraints
        if constraints:
            con = np.unique(np.concatenate([c.dofs for c in constraints]))
        else:
            con = np.zeros(0, dtype=np.int64)
        self.constrained = con
        self.free = np.setdiff1d(np.arange(size), con, assume_unique=True)
        self.size_reduced = self.free.size

    def lift(self, u_red, u_full):
        u = np.array(u_full, dtype=np.float64, copy=True)
        u[self.free] = u_red
        for c in self.constraints:
            u = c.lift(u)
        return u

    def lift_from_zeros(self, u_red):
        return self.lift(u_red, np.zeros(self.size))

    def lift_linear(self, v_red):
        """Derivative of lift w.r.t. u_red applied to v_red (Fixed values held
        constant, so they contribute zero)."""
        v = np.zeros(self.size)
        v[self.free] = v_red
        for c in self.constraints:
            if isinstance(c, Fixed):
                v[c.dofs] = 0.0
            else:
                v = c.lift(v)
        return v

    def reduce(self, u):
        return np.asarray(u)[self.free]

    def reduce_adjoint(self, r):
        r = np.array(r, dtype=np.float64, copy=True)
        for c in reversed(self.constraints):
            r = c.transpose(r)
        return r[self.free]

    def adapt_sparsity(self, S):
        for c in self.constraints:
            S = c.augment(S)
        return S[self.free][:, self.free]


# --------------------------------------------------------------------------
# Distance-2 coloring (sparse/_coloring.py)
# --------------------------------------------------------------------------


def pattern_from_element_dofs(element_dofs, size, diagonal_dofs=None):
    """pattern_from_compound's builder (sparse/_extraction.py:198-238): element
    dof pairs (repeat/tile), -1 entries filtered, diagonal dofs added,
    deduplicated through linearised indices."""
    ed = np.asarray(element_dofs, dtype=np.int64)
    rows_l, cols_l = [], []
    if ed.size:
        k = ed.shape[1]
        r = np.repeat(ed, k, axis=1).ravel()
        c = np.tile(ed, (1, k)).ravel()
        ok = (r >= 0) & (c >= 0)
        rows_l.append(r[ok])
        cols_l.append(c[ok])
    if diagonal_dofs is not None and np.size(diagonal_dofs):
        d = np.asarray(diagonal_dofs, dtype=np.int64).ravel()
        rows_l.append(d)
        cols_l.append(d)
    if not rows_l:
        return sps.csr_matrix((size, size), dtype=np.int8)
    lin = np.unique(np.concatenate(rows_l) * size + np.concatenate(cols_l))
    return sps.csr_matrix((np.ones(lin.size, dtype=np.int8), (lin // size, lin % size)),
                          shape=(size, size))


def distance2_adjacency(n, indptr, indices):
    """``get_distance2_adjacency`` (_coloring.py:27-48): rows of the boolean A @ A."""
    A = sps.csr_matrix((np.ones(len(indices), dtype=bool), indices, indptr), shape=(n, n))
    A2 = (A @ A).tocsr()
    A2.sort_indices()
    return [A2.indices[A2.indptr[i]:A2.indptr[i + 1]] for i in range(n)]


def greedy_coloring(adjacency):
    """``greedy_coloring`` (_coloring.py:136-153): natural order, smallest free color."""
    n = len(adjacency)
    colors = -np.ones(n, dtype=np.int32)
    for i in range(n):
        used = {colors[j] for j in adjacency[i] if colors[j] >= 0}
        c = 0
        while c in used:
            c += 1
        colors[i] = c
    return colors


def greedy_coloring_ldf(adjacency):
    """``greedy_coloring_ldf`` (_coloring.py:156-186): vertices in the order of
    ``np.argsort(degrees)[::-1]`` (degree = size of the A^2 row), each the
    smallest color its A^2 neighbours do not hold."""
    n = len(adjacency)
    degrees = np.array([len(a) for a in adjacency], dtype=np.int32)
    colors = -np.ones(n, dtype=np.int32)
    for i in np.argsort(degrees)[::-1]:
        used = {c for c in colors[adjacency[i]] if c >= 0}
        c = 0
        while c in used:
            c += 1
        colors[i] = c
    return colors


def largest_degree_first_distance2_colors(indptr, indices, n):
    """``largest_degree_first_distance2_colors`` (_coloring.py:286-300)."""
    return greedy_coloring_ldf(distance2_adjacency(n, np.asarray(indptr), np.asarray(indices)))


def is_distance2_coloring(indptr, indices, colors):
    """True when no row of the pattern has two columns of one color (what the
    colored decompression needs; tests/test_sparse.py:76-77 checks its result)."""
    for r in range(len(indptr) - 1):
        c = colors[indices[indptr[r]:indptr[r + 1]]]
        if np.unique(c).size != c.size:
            return False
    return True


def distance2_colors_py(indptr, indices, n):
    """``distance2_colors`` (_coloring.py:270-283), pure Python (small n only)."""
    return greedy_coloring(distance2_adjacency(n, np.asarray(indptr), np.asarray(indices)))


_here = os.path.dirname(os.path.abspath(__file__))
_CLIB = None


def _clib():
    global _CLIB
    if _CLIB is None:
        path = os.path.join(_here, "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle C library missing at {path}; run oracle/build.sh")
        lib = ctypes.CDLL(path)
        lib.oracle_distance2_greedy.restype = ctypes.c_int
        lib.oracle_distance2_greedy.argtypes = [
            ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.oracle_assemble.restype = ctypes.c_int
        lib.oracle_assemble.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
            ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
            ctypes.c_int]
        lib.oracle_max_threads.restype = ctypes.c_int
        _CLIB = lib
    return _CLIB


def distance2_colors(indptr, indices, n):
    """Same greedy as ``distance2_colors_py`` in plain C (oracle/coloring.c)."""
    ip = np.ascontiguousarray(indptr, dtype=np.int64)
    ix = np.ascontiguousarray(indices, dtype=np.int32)
    out = np.empty(n, dtype=np.int32)
    rc = _clib().oracle_distance2_greedy(n, ip.ctypes.data, ix.ctypes.data, out.ctypes.data)
    if rc < 0:
        raise RuntimeError("oracle coloring failed")
    return out


class ProblemOMP:
    """``Problem.residual`` / ``hvp`` / ``hessian_diag`` of a Tet4 or Hex8 mesh with
    linear elasticity or Neo-Hookean, in plain C over OpenMP threads
    (oracle/fem_omp.c: the same nodal element algebra, elements in parallel,
    atomic scatter into nodes).  For the CPU baseline at sizes where the NumPy
    Problem's per-element arrays do not fit; pinned against Problem in the
    oracle's tests."""

    _KIND = {"Tet4": 1, "Hex8": 2}

    def __init__(self, elem, coords, elements, density, threads=0):
        kind = self._KIND.get(type(elem).__name__)
        if kind is None or not isinstance(density, (LinearElasticity, NeoHookean)) or \
                np.ndim(density.mu) or np.ndim(density.lmbda):
            raise NotImplementedError("ProblemOMP: Tet4/Hex8 with LinearElasticity/NeoHookean, "
                                      "uniform parameters")
        self.kind = kind
        self.dens = 2 if isinstance(density, LinearElasticity) else 3
        self.mu, self.lmbda = float(density.mu), float(density.lmbda)
        self.X = np.ascontiguousarray(coords, dtype=np.float64)
        self.conn = np.ascontiguousarray(elements, dtype=np.int32)
        self.qp = np.ascontiguousarray(elem.quad_points, dtype=np.float64)
        self.qw = np.ascontiguousarray(elem.quad_weights, dtype=np.float64)
        self.n = self.X.shape[0] * 3
        self.threads = int(threads)

    def _run(self, mode, u, v=None):
        u = np.ascontiguousarray(u, dtype=np.float64)
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        out = np.empty(self.n)
        rc = _clib().oracle_assemble(
            self.kind, self.dens, self.mu, self.lmbda, self.X.ctypes.data, self.X.shape[0],
            self.conn.ctypes.data, self.conn.shape[0], self.qw.shape[0], self.qp.ctypes.data,
            self.qw.ctypes.data, u.ctypes.data, None if vv is None else vv.ctypes.data, mode,
            out.ctypes.data, self.threads)
        if rc != 0:
            raise RuntimeError("oracle_assemble failed")
        return out

    def residual(self, u):
        return self._run(0, u)

    def hvp(self, u, v):
        return self._run(1, u, v)

    def hessian_diag(self, u):
        return self._run(2, u)


def omp_max_threads() -> int:
    return int(_clib().oracle_max_threads())


# --------------------------------------------------------------------------
# Colored forward-mode Jacobian (sparse/base.py:108-270)
# --------------------------------------------------------------------------


def compute_rows_cols(indptr, indices, colors):
    """``compute_rows_cols`` (sparse/base.py:108-136)."""
    n = len(indptr) - 1
    rows = np.repeat(np.arange(n), np.diff(indptr))
    return rows, np.asarray(colors)[np.asarray(indices)]


def colored_jacfwd(jvp, indptr, indices, colors):
    """``jacfwd`` + ``colored_jacobian_batch`` (sparse/base.py:139-176,230-270).

    ``jvp(seed)`` returns J @ seed.  One JVP per color with seed = (colors == c),
    stacked into J_compressed (n, n_colors), then data = J_compressed[rows, col_colors].
    """
    colors = np.asarray(colors)
    n_colors = int(colors.max()) + 1
    cols = [jvp(np.where(colors == c, 1.0, 0.0)) for c in range(n_colors)]
    Jc = np.stack(cols, axis=0).T
    rows, col_colors = compute_rows_cols(indptr, indices, colors)
    return Jc[rows, col_colors]


# --------------------------------------------------------------------------
# Solvers (not in the reference; defined by the task: Jacobi-PCG + Newton)
# --------------------------------------------------------------------------


def pcg(A, b, x0=None, tol=1e-10, maxiter=100000, diag=None):
    """Jacobi-preconditioned CG on a SciPy CSR matrix; stops at ||r||/||b|| <= tol.

    Returns (x, iterations, relative residual of the recursive residual).
    """
    b = np.asarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    dinv = 1.0 / (A.diagonal() if diag is None else diag)
    r = b - A @ x
    bnorm = np.linalg.norm(b)
    if bnorm == 0.0:
        return x, 0, 0.0
    z = dinv * r
    p = z.copy()
    rz = r @ z
    it = 0
    rel = np.linalg.norm(r) / bnorm
    while rel > tol and it < maxiter:
        Ap = A @ p
        alpha = rz / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        z = dinv * r
        rz_new = r @ z
        it += 1
        rel = np.linalg.norm(r) / bnorm
        if rel <= tol:
            break
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, it, rel


def find_containing_polygons(points, polygons):
    """mesh.py:295-390: index of the first polygon holding each point (edges
    within 1e-8 count as inside; bounding-box reject; ray-casting parity), -1 if none."""
    pts = np.asarray(points, dtype=np.float64)
    out = np.full(pts.shape[0], -1, dtype=np.int64)
    for i, (px, py) in enumerate(pts):
        for e, V in enumerate(polygons):
            lo, hi = V.min(axis=0), V.max(axis=0)
            if not (lo[0] <= px <= hi[0] and lo[1] <= py <= hi[1]):
                continue
            W = np.roll(V, -1, axis=0)
            cross = (W[:, 0] - V[:, 0]) * (py - V[:, 1]) - (W[:, 1] - V[:, 1]) * (px - V[:, 0])
            on_seg = (np.minimum(V[:, 0], W[:, 0]) <= px) & (px <= np.maximum(V[:, 0], W[:, 0])) & \
                (np.minimum(V[:, 1], W[:, 1]) <= py) & (py <= np.maximum(V[:, 1], W[:, 1]))
            if np.any((np.abs(cross) <= 1e-8) & on_seg):
                out[i] = e
                break
            ycond = ((V[:, 1] <= py) & (W[:, 1] > py)) | ((W[:, 1] <= py) & (V[:, 1] > py))
            with np.errstate(divide="ignore", invalid="ignore"):
                xint = (W[:, 0] - V[:, 0]) * (py - V[:, 1]) / (W[:, 1] - V[:, 1]) + V[:, 0]
            if np.sum(ycond & (px < xint)) % 2 == 1:
                out[i] = e
                break
    return out


def find_containing_polygons_batched(points, polygons, chunk=64):
    """find_containing_polygons for meshes too large for the per-pair loop:
    the same bounding-box, edge (|cross| <= 1e-8) and ray-parity tests and the
    same first-polygon rule, evaluated for chunks of points against all
    polygons at once with numpy broadcasting."""
    pts = np.asarray(points, dtype=np.float64)
    V = np.asarray(polygons, dtype=np.float64)
    W = np.roll(V, -1, axis=1)
    lo, hi = V.min(axis=1), V.max(axis=1)                     # (E, 2)
    V0, V1, W0, W1 = V[None, :, :, 0], V[None, :, :, 1], W[None, :, :, 0], W[None, :, :, 1]
    out = np.full(pts.shape[0], -1, dtype=np.int64)
    for a in range(0, pts.shape[0], chunk):
        px = pts[a:a + chunk, 0][:, None]
        py = pts[a:a + chunk, 1][:, None]
        inbox = (lo[None, :, 0] <= px) & (px <= hi[None, :, 0]) & \
            (lo[None, :, 1] <= py) & (py <= hi[None, :, 1])   # (c, E)
        PX, PY = px[:, :, None], py[:, :, None]
        cross = (W0 - V0) * (PY - V1) - (W1 - V1) * (PX - V0)
        on_seg = (np.minimum(V0, W0) <= PX) & (PX <= np.maximum(V0, W0)) & \
            (np.minimum(V1, W1) <= PY) & (PY <= np.maximum(V1, W1))
        boundary = np.any((np.abs(cross) <= 1e-8) & on_seg, axis=2)
        ycond = ((V1 <= PY) & (W1 > PY)) | ((W1 <= PY) & (V1 > PY))
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = (W0 - V0) * (PY - V1) / (W1 - V1) + V0
        parity = np.sum(ycond & (PX < xint), axis=2) % 2 == 1
        inside = inbox & (boundary | parity)
        out[a:a + chunk] = np.where(inside.any(axis=1), np.argmax(inside, axis=1), -1)
    return out


def interpolate(elem, coords, elements, nodal, points, batched=False):
    """Operator.interpolate (operator.py:399-468): one Newton step from the
    first quadrature point in the containing element, NaN outside
    (``batched``: locate with find_containing_polygons_batched)."""
    X = np.asarray(coords)
    el = np.asarray(elements)
    locate = find_containing_polygons_batched if batched else find_containing_polygons
    idx = locate(points, X[el])
    nodal = np.asarray(nodal, dtype=np.float64)
    out = np.full((len(points),) + nodal.shape[1:], np.nan)
    xi0 = elem.quad_points[0]
    for i, e in enumerate(idx):
        if e < 0:
            continue
        Xe = X[el[e]]
        x0 = elem.N(xi0) @ Xe
        J = (elem.dN(xi0) @ Xe).T                      # dx_i / dxi_k
        xi = xi0 + np.linalg.solve(J, np.asarray(points[i]) - x0)
        out[i] = np.tensordot(elem.N(xi), nodal[el[e]], axes=1)
    return out


def project(elem, coords, elements, field_q, lifter=None):
    """L2 projection (operator.py:523-554, utils.py:118-262) with the scalar
    mass matrix and one right-hand side per component: M_ab = sum N_a N_b w,
    b_a = sum N_a f w.  With a lifter (utils.py:225-246) the test functions sit
    at the free dofs and the trial field is lifted, so the matrix is the
    Jacobian of the lifted form, E_free^T M P with P the linear part of the lift
    (a Periodic slave's column added to its master's; for Fixed constraints only
    it is M[free][:, free]); the solution is lifted from zeros.  Dense assembly;
    for small meshes."""
    w = integration_weights(elem, coords, elements)                 # (E, Q)
    Nq = np.stack([elem.N(xi) for xi in elem.quad_points])         # (Q, nne)
    n = coords.shape[0]
    M = np.zeros((n, n))
    Me = np.einsum("eq,qa,qb->eab", w, Nq, Nq)
    for e, el in enumerate(np.asarray(elements)):
        M[np.ix_(el, el)] += Me[e]
    f = np.asarray(field_q, dtype=np.float64)
    vs = f.shape[2:]
    f2 = f.reshape(f.shape[0], f.shape[1], -1)
    b = np.zeros((n, f2.shape[2]))
    np.add.at(b, np.asarray(elements), np.einsum("eq,qa,eqc->eac", w, Nq, f2))
    if lifter is None:
        x = np.linalg.solve(M, b)
    else:
        fr = lifter.free
        P = np.stack([lifter.lift_linear(e) for e in np.eye(fr.size)], axis=1)   # (n, n_free)
        xr = np.linalg.solve(M[fr] @ P, b[fr])
        x = np.stack([lifter.lift_from_zeros(xr[:, c]) for c in range(xr.shape[1])], axis=1)
    return x.reshape((n,) + vs)


def pcg_single_reduction(A, b, x0=None, tol=1e-10, maxiter=100000, diag=None):
    """Jacobi-preconditioned CG in the single-reduction form of Chronopoulos and
    Gear (J. Comput. Appl. Math. 25, 1989): u = D^-1 r, w = A u,
    gamma = r.u, delta = w.u, beta = gamma / gamma_old,
    alpha = gamma / (delta - beta gamma / alpha_old), p = u + beta p,
    s = w + beta s, x += alpha p, r -= alpha s.  Equal to ``pcg`` in exact
    arithmetic; the convergence test is on r before each update.

    Returns (x, iterations, relative residual of the recursive residual).
    """
    b = np.asarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    dinv = 1.0 / (A.diagonal() if diag is None else diag)
    r = b - A @ x
    bnorm = np.linalg.norm(b)
    if bnorm == 0.0:
        return x, 0, 0.0
    u = dinv * r
    gamma = r @ u
    it = 0
    p = s = None
    alpha_old = gamma_old = None
    while True:
        rel = np.linalg.norm(r) / bnorm
        if rel <= tol or it >= maxiter:
            return x, it, rel
        w = A @ u
        delta = w @ u
        if it == 0:
            beta, denom = 0.0, delta
            p, s = u.copy(), w.copy()
        else:
            beta = gamma / gamma_old
            denom = delta - beta * gamma / alpha_old
            p = u + beta * p
            s = w + beta * s
        alpha = gamma / denom
        x += alpha * p
        r = r - alpha * s
        u = dinv * r
        alpha_old, gamma_old = alpha, gamma
        gamma = r @ u
        it += 1


def hessian_csr(prob: Problem, lifter: Lifter, u_full, pattern, colors):
    """Colored sparse Hessian of E_free(u) = E(lift(u)) on the reduced pattern."""
    def jvp(seed):
        return lifter.reduce_adjoint(prob.hvp(u_full, lifter.lift_linear(seed)))
    data = colored_jacfwd(jvp, pattern.indptr, pattern.indices, colors)
    return sps.csr_matrix((data, pattern.indices, pattern.indptr), shape=pattern.shape)


def newton(prob: Problem, lifter: Lifter, u_full, pattern, colors, tol=1e-10,
           lin_tol=1e-10, max_newton=20, maxiter=100000):
    """Newton on the free dofs with a colored-CSR Hessian and Jacobi-PCG.

    Returns (u_full, newton_iterations, total_cg_iterations, residual history).
    """
    u = np.array(u_full, dtype=np.float64, copy=True)
    hist = []
    cg_total = 0
    r0 = None
    for it in range(max_newton):
        r = lifter.reduce_adjoint(prob.residual(u))
        rn = np.linalg.norm(r)
        hist.append(rn)
        if r0 is None:
            r0 = rn
        if rn <= tol * max(r0, 1e-300):
            return u, it, cg_total, hist
        K = hessian_csr(prob, lifter, u, pattern, colors)
        du, k, _ = pcg(K, -r, tol=lin_tol, maxiter=maxiter)
        cg_total += k
        u = lifter.lift(lifter.reduce(u) + du, u)
    r = lifter.reduce_adjoint(prob.residual(u))
    hist.append(np.linalg.norm(r))
    return u, max_newton, cg_total, hist


def face_traction_loads(coords, n, axis=0, at=1.0, t=(1e-3, 0.0, 0.0), h=None):
    """Consistent nodal loads of a uniform traction on a structured Hex8 face.

    Each face node carries t * h^2 * {1/4 corner, 1/2 edge, 1 interior}.
    """
    m = n + 1
    h = 1.0 / n if h is None else h
    nodes = np.where(np.isclose(coords[:, axis], at))[0]
    other = [a for a in range(3) if a != axis]
    f = np.zeros(coords.shape[0] * 3)
    for nd in nodes:
        w = 1.0
        for a in other:
            idx = int(round(coords[nd, a] / h))
            if idx == 0 or idx == m - 1:
                w *= 0.5
        for c in range(3):
            f[nd * 3 + c] += t[c] * h * h * w
    return f
