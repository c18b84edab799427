"""MI355X-native (gfx950) integrated surface/subsurface flow solver.

Hot path: hand-written HIP kernels behind the C-ABI of include/dhara_hip.h
(libdhara_hip.so in this directory).  `FlowModel` is the Python host mirror;
`inputs` builds the synthetic workloads.
"""
from . import _abi
from ._abi import (BC_DIRICHLET, BC_FLUX, FP64, INIT_FROM_PSI, INIT_FROM_THETA, MIXED,
                   PICARD_ABS_PSI, PICARD_REL_PSI, PICARD_REL_THETA, THETA_CARRY_CONSISTENT,
                   THETA_CARRY_REFERENCE, TOP_PRESCRIBED, TOP_SURFACE, Grid, Params, StepStats,
                   default_params)
from .model import DharaError, Ensemble, FlowModel, device_info, load_hip_library

__all__ = [
    "Grid", "Params", "StepStats", "default_params", "FlowModel", "Ensemble", "DharaError", "device_info",
    "load_hip_library", "FP64", "MIXED", "PICARD_ABS_PSI", "PICARD_REL_PSI", "PICARD_REL_THETA",
    "THETA_CARRY_CONSISTENT", "THETA_CARRY_REFERENCE", "TOP_SURFACE", "TOP_PRESCRIBED",
    "INIT_FROM_PSI", "INIT_FROM_THETA", "BC_DIRICHLET", "BC_FLUX", "_abi",
]
