"""weldform_amd: MI355X-native explicit SPH step of WeldForm.

The compute path is libweldform_hip.so (HIP kernels for gfx950 behind the
C-ABI in include/weldform_hip.h).  This package holds its build recipe,
the ctypes binding and the workload builders.
"""
from . import capi, cases  # noqa: F401
from .capi import Solver, WFError, gen_cylinder  # noqa: F401
