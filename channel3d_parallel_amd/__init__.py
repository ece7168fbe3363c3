"""MI355X-native (gfx950) channel3D fractional-step projection path.

The compute path is the HIP library ``libc3d.so`` behind the C ABI of
``include/c3d.h``; :mod:`channel3d_parallel_amd.solver` mirrors the
reference's module procedures on top of it and the Fortran host modules in
``fortran/`` bind the same entry points with ``iso_c_binding``.
"""
from . import fields  # noqa: F401
from ._lib import LIB_PATH, load  # noqa: F401
from .solver import C3DError, Channel  # noqa: F401

__all__ = ["Channel", "C3DError", "load", "LIB_PATH", "fields"]
