"""MI355X-native 2-D poroelastic (Biot) C-PML solver: Python handle over libpoe2d.so.

The compute path is the HIP library built from poe2d_amd/csrc; importing this package does not
load it, the first Solver does (and raises if it is missing).
"""
from . import _abi
from ._abi import (F32, F64, P, SRC_FORCE_Y_RICKER, SRC_PRESSURE_DGAUSS, VX, VY, Poe2dError,
                   load_library)
from .solver import Problem, ReceiverSpec, Solver, SourceSpec

__all__ = ["F32", "F64", "P", "VX", "VY", "SRC_FORCE_Y_RICKER", "SRC_PRESSURE_DGAUSS",
           "Poe2dError", "load_library", "Problem", "ReceiverSpec", "Solver", "SourceSpec"]
