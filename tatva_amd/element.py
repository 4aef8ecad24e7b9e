"""Reference elements: Tri3, Tetrahedron4, Hexahedron8.

Mirrors ``tatva.element`` (element/base.py:26-141,245-265,448-568): each element
carries its quadrature rule (default or user-supplied ``quad_points``/
``quad_weights``), reference nodes and shape-function tables.  All per-element
arithmetic (Jacobians, gradients, interpolation) runs in the gfx950 kernels of
``Operator``; the host-side methods here only describe the element.
"""

from __future__ import annotations

import numpy as np

from . import _native as nat


class Element:
    """Base class (element/base.py:26-68).  Immutable; compared by type and quadrature."""

    kind: int
    dim: int
    nne: int

    def __init__(self, quad_points=None, quad_weights=None):
        if quad_points is not None and quad_weights is not None:
            qp = np.asarray(quad_points, dtype=np.float64).reshape(-1, self.dim)
            qw = np.asarray(quad_weights, dtype=np.float64).reshape(-1)
            if qp.shape[0] != qw.shape[0]:
                raise ValueError("quad_points and quad_weights must have the same length")
            if not 1 <= qw.shape[0] <= 27:
                raise ValueError("between 1 and 27 quadrature points are supported")
            self.quad_points, self.quad_weights = qp, qw
        else:
            self.quad_points, self.quad_weights = self._default_quadrature()

    def __eq__(self, other) -> bool:
        return (type(self) is type(other)
                and np.array_equal(self.quad_points, other.quad_points)
                and np.array_equal(self.quad_weights, other.quad_weights))

    def __hash__(self) -> int:
        return hash((type(self), self.quad_points.tobytes(), self.quad_weights.tobytes()))

    def __repr__(self) -> str:
        return f"{type(self).__name__}(n_quad={self.quad_weights.shape[0]})"

    def _reference_nodes(self) -> np.ndarray:
        raise NotImplementedError

    def _default_quadrature(self):
        raise NotImplementedError


class Tri3(Element):
    """3-node linear triangle (element/base.py:245-265)."""

    kind, dim, nne = nat.TRI3, 2, 3

    def _reference_nodes(self):
        return np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])

    def _default_quadrature(self):
        return np.array([[1.0 / 3, 1.0 / 3]]), np.array([1.0 / 2])


class Tetrahedron4(Element):
    """4-node linear tetrahedron (element/base.py:448-472)."""

    kind, dim, nne = nat.TET4, 3, 4

    def _reference_nodes(self):
        return np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])

    def _default_quadrature(self):
        return np.array([[0.25, 0.25, 0.25]]), np.array([1.0 / 6])


class Hexahedron8(Element):
    """8-node trilinear hexahedron with 2x2x2 Gauss rule (element/base.py:475-568)."""

    kind, dim, nne = nat.HEX8, 3, 8

    def _reference_nodes(self):
        return np.array([[-1.0, -1.0, -1.0], [1.0, -1.0, -1.0], [1.0, 1.0, -1.0],
                         [-1.0, 1.0, -1.0], [-1.0, -1.0, 1.0], [1.0, -1.0, 1.0],
                         [1.0, 1.0, 1.0], [-1.0, 1.0, 1.0]])

    def _default_quadrature(self):
        a = 1.0 / np.sqrt(3.0)
        pts = a * np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                            [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
        return pts, np.ones(8)


__all__ = ["Element", "Tri3", "Tetrahedron4", "Hexahedron8"]
