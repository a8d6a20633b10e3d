"""pyfantom_amd -- MI355X (gfx950) implementation of the structured SIMP
finite-element solve path, behind the object API of pyFANTOM's CUDA backend.

>>> from pyfantom_amd import (StructuredMesh3D, StructuredStiffnessKernel, MultiGrid,
...                           FiniteElement, StructuredFilter3D, MinimumCompliance, OC, MMA)

The names match ``pyFANTOM.CUDA`` (pyFANTOM/CUDA/__init__.py:10-16) for the
structured path.  Device arrays are torch tensors on the HIP device; every
hot operation is a kernel of ``libfantom_hip.so`` called through its C ABI
(``include/fantom.h``).
"""
from .geom import StructuredFilter2D, StructuredFilter3D, StructuredMesh2D, StructuredMesh3D  # noqa: F401
from .physics import LinearElasticity  # noqa: F401
from .stiffness import StructuredStiffnessKernel  # noqa: F401
from .solvers import CG, MultiGrid  # noqa: F401
from .fe import FiniteElement  # noqa: F401
from .problem import MinimumCompliance  # noqa: F401
from .optimizers import MMA, OC  # noqa: F401
from . import checkpoint  # noqa: F401

__all__ = ["StructuredMesh2D", "StructuredMesh3D", "StructuredFilter2D", "StructuredFilter3D", "LinearElasticity",
           "StructuredStiffnessKernel", "MultiGrid", "CG", "FiniteElement", "MinimumCompliance", "OC", "MMA"]
