"""CPU FP64 oracle for the channel3D projection path (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use this package, to test and time against.  The product package
``channel3d_parallel_amd`` never imports it.

The C library ``liborc.so`` (built by ``oracle/Makefile``) restates the
reference's Fortran routines line by line; this module wraps it with numpy
arrays in the reference's Fortran (column-major, ghost-inclusive) layout.
"""
from .orc import (  # noqa: F401
    Grid,
    lib,
    build,
)
