"""tatva_amd: MI355X-native (gfx950 HIP) element-assembly, colored-Hessian,
matrix-free-HVP and Jacobi-PCG path with tatva's API.

Drop-in names (tatva v0.11.1): ``Mesh``, ``Operator``, ``element.Tri3 /
Tetrahedron4 / Hexahedron8``, ``sparse.pattern_from_mesh / ColoredMatrix /
jacfwd / linearized_jacfwd``, ``lifter.Lifter / Fixed / RuntimeValue``.
Energies are built-in densities (``energies``); ``grad``/``jvp``/``linearize``
stand in for the corresponding JAX transformations of those energies.
"""

from . import element, energies, lifter, solvers, sparse  # noqa: F401
from .element import Hexahedron8, Tetrahedron4, Tri3  # noqa: F401
from .energies import (LinearElasticity, NeoHookean, Poisson, grad, jvp,  # noqa: F401
                       linearize)
from .mesh import Mesh  # noqa: F401
from .operator import Operator  # noqa: F401

__version__ = "0.1.0"

__all__ = ["Mesh", "Operator", "element", "energies", "lifter", "solvers", "sparse",
           "Tri3", "Tetrahedron4", "Hexahedron8", "Poisson", "LinearElasticity", "NeoHookean",
           "grad", "jvp", "linearize"]
