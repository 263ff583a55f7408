"""pscf_amd: MI355X (gfx950) HIP implementation of PSCF's SCFT inner loop.

The compute path is libpscf_hip.so (pscf_amd/csrc, C ABI in include/pscf_hip.h);
the Python modules mirror the reference's Fortran module interfaces.
"""
from .system import System, diblock  # noqa: F401
