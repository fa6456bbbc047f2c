"""treesap_amd: MI355X-native (gfx950) explicit RK4 time stepping for TreESaP models.

The compute path is the HIP library ``libtreesap_hip.so`` (C ABI in
include/treesap_hip.h); this package holds its ctypes binding, the host-side
solver object and the seeded workload definitions.
"""
from . import _lib
from .solver import Batch, Level, Solver
from .workloads import Case, make_desc

__all__ = ["Batch", "Level", "Solver", "Case", "make_desc", "_lib"]
__version__ = "0.1.0"
