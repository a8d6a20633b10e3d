"""pyfantom_amd -- MI355X (gfx950) implementation of the structured SIMP
finite-element solve path, behind the object API of pyFANTOM's CUDA backend.

>>> from pyfantom_amd import (StructuredMesh3D, StructuredStiffnessKernel, MultiGrid,
...                           FiniteElement, StructuredFilter3D, MinimumCompliance, OC, MMA)

The names match ``pyFANTOM.CUDA`` (pyFANTOM/CUDA/__init__.py:10-16) for the
structured path, plus ``GeneralMesh`` with ``UniformStiffnessKernel`` /
``GeneralStiffnessKernel`` (unstructured triangles, quads, tetrahedra, hexes)
solved by ``CG``.  Device arrays are torch tensors on the HIP device; every
hot operation is a kernel of ``libfantom_hip.so`` called through its C ABI
(``include/fantom.h``).
"""
from .geom import (GeneralMesh, StructuredFilter2D, StructuredFilter3D, StructuredMesh2D,  # noqa: F401
                   StructuredMesh3D)
from .physics import LinearElasticity  # noqa: F401
from .stiffness import GeneralStiffnessKernel, StructuredStiffnessKernel, UniformStiffnessKernel  # noqa: F401
from .solvers import CG, MultiGrid  # noqa: F401
from .fe import FiniteElement  # noqa: F401
from .problem import MinimumCompliance  # noqa: F401
from .optimizers import MMA, OC  # noqa: F401
from . import checkpoint  # noqa: F401

__all__ = ["StructuredMesh2D", "StructuredMesh3D", "GeneralMesh", "StructuredFilter2D", "StructuredFilter3D",
           "LinearElasticity", "StructuredStiffnessKernel", "UniformStiffnessKernel", "GeneralStiffnessKernel",
           "MultiGrid", "CG", "FiniteElement", "MinimumCompliance", "OC", "MMA"]
