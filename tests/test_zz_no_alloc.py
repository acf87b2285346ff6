"""pv_evaluate_device_noalloc (the entry point that allocates nothing) against pv_evaluate.

The test allocates its own device buffers through the HIP runtime (the entry point's callers
own theirs); the entry point itself only uses its fixed per-GPU scratch array.  Two calls in a
row, the second with a smaller population, must each give pv_evaluate's velocities bit for bit
when both use the built-in launch shape (a handle with an empty launch table), and stay
within the FP64 bar (normwise 1e-12) of the tuned handle; mixed precision within 1e-6 of
FP64.  Errors come back with pv_evaluate's messages, and a call after an error still works.
"""
import ctypes

import numpy as np
import pytest

from pointvortex_amd import (BackendRuntimeError, HipKernel, InvalidArgument, SimParams,
                             VortexSystem)
from pointvortex_amd import backend
from pointvortex_amd import workloads as wl
from tests.conftest import normwise

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    backend.load_library()          # loads the HIP runtime the library links
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            lib = ctypes.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.skip("HIP runtime library not found")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return lib


class Buffers:
    """The caller's device arrays x, y, circulation, u, v of n doubles each."""

    def __init__(self, hip, n):
        self.hip, self.n, self.ptr = hip, n, []
        for _ in range(5):
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), max(1, n) * 8) == 0
            self.ptr.append(p.value)

    def upload(self, x, y, g):
        for p, a in zip(self.ptr[:3], (x, y, g)):
            a = np.ascontiguousarray(a, dtype=np.float64)
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0

    def velocities(self):
        out = []
        for p in self.ptr[3:]:
            a = np.empty(self.n)
            assert self.hip.hipMemcpy(a.ctypes.data, p, a.nbytes, D2H) == 0
            out.append(a)
        return out

    def free(self):
        for p in self.ptr:
            self.hip.hipFree(p)


def population(geometry, n, seed):
    if geometry == "infinite":
        return wl.infinite(n, seed), SimParams(boundaryCondition="infinite")
    if geometry == "periodic":
        return (wl.periodic(n, seed, box_length=2.0),
                SimParams(boundaryCondition="periodic", boxLengthX=2.0, boxLengthY=2.0))
    return wl.disk(n, seed), SimParams(boundaryCondition="disk")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
def test_noalloc_matches_pv_evaluate_twice(geometry, precision):
    hip = _hip()
    for n, seed in ((5000, 1), (778, 2)):          # the second call is smaller
        pop, params = population(geometry, n, seed)
        params.precision = precision
        buf = Buffers(hip, n)
        try:
            buf.upload(pop.x, pop.y, pop.circulation)
            # Scratch contents of the previous call (or test) must not leak into this one.
            backend.evaluate_device_noalloc(params, *buf.ptr[:3], n, *buf.ptr[3:])
            u, v = buf.velocities()
        finally:
            buf.free()
        plain = HipKernel(params)
        plain.set_launch_table([])                  # the built-in shape, as the entry point
        ref = plain.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
        assert np.array_equal(u, ref.x) and np.array_equal(v, ref.y), (n, geometry, precision)
        tuned = HipKernel(SimParams(**{**params.__dict__, "precision": "fp64"}))
        r64 = tuned.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
        bar = 1e-12 if precision == "fp64" else 1e-6
        assert normwise(u, v, r64.x, r64.y) <= bar, (n, geometry, precision)


def test_noalloc_errors_and_limits():
    hip = _hip()
    params = SimParams(boundaryCondition="infinite")
    with pytest.raises(InvalidArgument, match="no-allocation limit"):
        backend.evaluate_device_noalloc(params, 0, 0, 0, backend.NOALLOC_MAX_VORTICES + 1, 0, 0)
    pop = wl.infinite(300, 4)
    x = pop.x.copy()
    x[7], x[8] = x[3], x[3]
    y = pop.y.copy()
    y[7], y[8] = y[3], y[3]
    buf = Buffers(hip, 300)
    try:
        buf.upload(x, y, pop.circulation)
        with pytest.raises(BackendRuntimeError,
                           match="coincident vortices in singular point-vortex kernel"):
            backend.evaluate_device_noalloc(params, *buf.ptr[:3], 300, *buf.ptr[3:])
        g = pop.circulation.copy()
        g[5] = np.nan
        buf.upload(pop.x, pop.y, g)
        with pytest.raises(InvalidArgument, match="vortex data must be finite"):
            backend.evaluate_device_noalloc(params, *buf.ptr[:3], 300, *buf.ptr[3:])
        per = SimParams(boundaryCondition="periodic")
        buf.upload(pop.x, pop.y, np.abs(pop.circulation))
        with pytest.raises(InvalidArgument, match="zero total circulation"):
            backend.evaluate_device_noalloc(per, *buf.ptr[:3], 300, *buf.ptr[3:])
        # After the errors the same scratch serves a good call.
        buf.upload(pop.x, pop.y, pop.circulation)
        backend.evaluate_device_noalloc(params, *buf.ptr[:3], 300, *buf.ptr[3:])
        u, v = buf.velocities()
    finally:
        buf.free()
    plain = HipKernel(params)
    plain.set_launch_table([])
    ref = plain.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
    assert np.array_equal(u, ref.x) and np.array_equal(v, ref.y)
