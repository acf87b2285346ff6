"""point_vortex_hip: an MI355X (gfx950) velocity backend for PointVortex.

The compute path is libpvhip.so (hand-written HIP kernels behind the C-ABI of
include/pv_hip.h); this package is the host-side mirror of the reference's backend interface.
"""
from .backend import (  # noqa: F401
    BackendRuntimeError,
    HipKernel,
    Invariants,
    InvalidArgument,
    LogicError,
    OutOfRange,
    SimParams,
    StepResult,
    VectorField,
    VortexSystem,
    backendName,
    backendRuntimeDetails,
    computeInvariants,
    declared_symbols,
    load_library,
    makeBackendKernel,
)
