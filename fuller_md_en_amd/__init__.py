"""fuller_md_en_amd -- MI355X-native force engine for fullerene-crystal NPT MD.

The compute path is libfmd.so: hand-written gfx950 HIP kernels behind the C
ABI of include/fmd.h.  This package is the Python face of that ABI: the
reference's force / list / integrator functions with the same arguments
(``rigid``), and the HBM-resident engines used by the drivers and bench.py.
"""
from ._native import FmdError, lib  # noqa: F401

__all__ = ["FmdError", "lib"]
