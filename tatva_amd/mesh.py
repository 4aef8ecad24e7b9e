"""Mesh: node coordinates and element connectivity (tatva/mesh.py:53-231).

The mesh is host-side setup data (FP64 coordinates, int32 connectivity, the
layouts tatva uses); ``Operator`` uploads it once and keeps it resident on the
GPU together with the node->element incidence used by the assembly kernels.
"""

from __future__ import annotations

from dataclasses import dataclass, replace
from enum import Enum

import numpy as np


class ElementType(Enum):
    """tatva/mesh.py:31-38."""

    TRIANGLE = "triangle"
    QUAD = "quad"
    TETRAHEDRON = "tetrahedron"
    HEXAHEDRON = "hexahedron"


def _as_np(a, dtype):
    try:
        import torch  # noqa: PLC0415
        if isinstance(a, torch.Tensor):
            a = a.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return np.ascontiguousarray(np.asarray(a), dtype=dtype)


def _min_corner_volume(coords, elements, kind) -> float:
    """Smallest signed volume (times 6) of the corner tetrahedra of the elements:
    each Tet4 itself; for Hex8 the tet spanned by the three edges at each of the
    eight corners (its sign is that of det J at the corner)."""
    X = coords[elements]
    if kind is ElementType.TETRAHEDRON:
        corners = [(0, 1, 2, 3)]
    else:
        # corner, then its neighbours along +xi, +eta, +zeta (Hexahedron8 order)
        corners = [(0, 1, 3, 4), (1, 2, 0, 5), (2, 3, 1, 6), (3, 0, 2, 7),
                   (4, 7, 5, 0), (5, 4, 6, 1), (6, 5, 7, 2), (7, 6, 4, 3)]
    vmin = np.inf
    for c, a, b, d in corners:
        e1, e2, e3 = X[:, a] - X[:, c], X[:, b] - X[:, c], X[:, d] - X[:, c]
        vmin = min(vmin, float(np.einsum("ij,ij->i", e1, np.cross(e2, e3)).min()))
    return vmin


@dataclass(frozen=True, eq=False)
class Mesh:
    """``Mesh(coords, elements)``: coords (n_nodes, n_dim), elements (n_elements, nne)."""

    coords: np.ndarray
    elements: np.ndarray

    def __post_init__(self):
        c = np.asarray(self.coords)
        if c.dtype.kind not in "fiu":
            raise TypeError("Mesh coordinates must be numeric.")
        e = np.asarray(self.elements)
        if e.dtype.kind not in "iu":
            raise TypeError("Mesh element connectivity must contain integer indices.")
        object.__setattr__(self, "coords", _as_np(self.coords, np.float64))
        object.__setattr__(self, "elements", _as_np(self.elements, np.int32))

    def _replace(self, **changes) -> "Mesh":
        return replace(self, **changes)

    def set_coords(self, new_coords) -> "Mesh":
        """A mesh with the same connectivity and new coordinates (mesh.py:77-85)."""
        return replace(self, coords=new_coords)

    @property
    def n_nodes(self) -> int:
        return int(self.coords.shape[0])

    @property
    def n_elements(self) -> int:
        return int(self.elements.shape[0])

    # -------------------------------------------------------- generators ----

    @classmethod
    def unit_square(cls, n_x: int, n_y: int, *, type="triangle", dim: int = 2) -> "Mesh":
        """``Mesh.unit_square`` (mesh.py:146-156)."""
        return cls.rectangle((0.0, 1.0), (0.0, 1.0), n_x, n_y, type=type, dim=dim)

    @classmethod
    def rectangle(cls, x, y, n_x: int, n_y: int, *, type="triangle", dim: int = 2) -> "Mesh":
        """Structured triangle mesh of [x0,x1] x [y0,y1] (mesh.py:158-205).

        Node (i, j) has id ``i*(n_y+1)+j``; each cell is split into
        ``[n0, n1, n3]`` and ``[n0, n3, n2]`` with n0=(i,j), n1=(i+1,j),
        n2=(i,j+1), n3=(i+1,j+1), cells ordered i-major.
        """
        kind = ElementType(type)
        if kind is not ElementType.TRIANGLE:
            raise NotImplementedError(f"Element type {type} not implemented.")
        xs = np.linspace(x[0], x[1], n_x + 1)
        ys = np.linspace(y[0], y[1], n_y + 1)
        coords = np.stack(np.meshgrid(xs, ys, indexing="ij"), axis=-1).reshape(-1, 2)
        i, j = np.meshgrid(np.arange(n_x), np.arange(n_y), indexing="ij")
        n0 = (i * (n_y + 1) + j).ravel()
        n1 = n0 + (n_y + 1)
        n2 = n0 + 1
        n3 = n1 + 1
        tris = np.stack([np.stack([n0, n1, n3], -1), np.stack([n0, n3, n2], -1)], axis=1)
        elements = tris.reshape(-1, 3).astype(np.int32)
        if dim == 3:
            coords = np.hstack([coords, np.zeros((coords.shape[0], 1))])
        return cls(coords=coords, elements=elements)

    @classmethod
    def box(cls, n, *, type="hexahedron", lengths=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0),
            jitter: float = 0.0, seed: int = 0) -> "Mesh":
        """Structured 3-D box mesh with ``n = (nx, ny, nz)`` cells (or one int).

        Nodes are numbered ``i + (nx+1) j + (nx+1)(ny+1) k``.  Hexahedra use the
        ``Hexahedron8`` node order (element/base.py:478-491); cells are ordered
        with i fastest.  Tetrahedra split each cell into the six tets of
        ``tests/test_sparse_tracer.py:47-68`` (all sharing the n0-n7 diagonal).
        ``jitter > 0`` moves every interior node by U(-jitter*h, jitter*h) per
        coordinate, drawn as ``default_rng(seed).uniform(-jitter, jitter, (N, 3))``.
        """
        nx, ny, nz = (n, n, n) if np.isscalar(n) else tuple(int(v) for v in n)
        L = np.asarray(lengths, dtype=np.float64)
        o = np.asarray(origin, dtype=np.float64)
        xs = o[0] + np.linspace(0.0, L[0], nx + 1)
        ys = o[1] + np.linspace(0.0, L[1], ny + 1)
        zs = o[2] + np.linspace(0.0, L[2], nz + 1)
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=-1)
        sx, sy, sz = 1, nx + 1, (nx + 1) * (ny + 1)
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        c0 = (i * sx + j * sy + k * sz).ravel()
        # cell corners: (x, y, z) offsets
        v = {(a, b, c): c0 + a * sx + b * sy + c * sz for a in (0, 1) for b in (0, 1) for c in (0, 1)}
        kind = ElementType(type)
        if kind is ElementType.HEXAHEDRON:
            order = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                     (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
            elements = np.stack([v[t] for t in order], axis=-1)
        elif kind is ElementType.TETRAHEDRON:
            n0, n1, n2, n3 = v[0, 0, 0], v[1, 0, 0], v[0, 1, 0], v[1, 1, 0]
            n4, n5, n6, n7 = v[0, 0, 1], v[1, 0, 1], v[0, 1, 1], v[1, 1, 1]
            tets = [(n0, n1, n3, n7), (n0, n1, n7, n5), (n0, n5, n7, n4),
                    (n0, n3, n2, n7), (n0, n2, n6, n7), (n0, n6, n4, n7)]
            elements = np.stack([np.stack(t, axis=-1) for t in tets], axis=1).reshape(-1, 4)
        else:
            raise NotImplementedError(f"Element type {type} not implemented.")
        if jitter > 0.0:
            rng = np.random.default_rng(seed)
            h = L / np.array([nx, ny, nz], dtype=np.float64)
            delta = rng.uniform(-jitter, jitter, size=coords.shape) * h
            lo = np.isclose(coords, o)
            hi = np.isclose(coords, o + L)
            interior = ~(lo | hi).any(axis=1)
            coords = coords + delta * interior[:, None]
            if _min_corner_volume(coords, elements, kind) <= 0.0:
                raise ValueError(f"jitter={jitter} inverts an element; use a smaller jitter")
        return cls(coords=coords, elements=elements.astype(np.int32))

    @classmethod
    def unit_cube(cls, n: int, *, type="hexahedron", jitter: float = 0.0, seed: int = 0) -> "Mesh":
        """[0, 1]^3 with n cells per side (see ``box``)."""
        return cls.box(n, type=type, jitter=jitter, seed=seed)


__all__ = ["Mesh", "ElementType"]
