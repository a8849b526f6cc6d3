"""fastgpr -- MI355X-native 2-D TMz FDTD forward / adjoint for GPR FWI.

Drop-in for the reference's ``src.compute`` surface::

    from fastgpr.compute import compute          # was: from src.compute import compute

The solver kernels live in ``libfastgpr.so`` (HIP, gfx950), loaded through
ctypes on import; importing fails if the library has not been built.
"""
from . import _native  # noqa: F401  (loads libfastgpr.so or raises ImportError)
from .compute import GPRFWI, GPRFWIModel, compute, initialization  # noqa: F401
from .solver import release_graphs  # noqa: F401

__all__ = ["compute", "GPRFWI", "GPRFWIModel", "initialization", "release_graphs"]
