"""Python host mirror of the reference's backend interface over the libpvhip C-ABI.

Names and argument meanings follow src/compute.h:18-36 (VelocityKernel), src/backend.h:9-16
(makeBackendKernel & co.), src/vortex.h (VortexSystem, VectorField), src/compute.h:11-16
(StepResult) and src/compute.h:78-81 (Invariants).  Exceptions mirror the reference's C++
exception types:

    std::invalid_argument -> InvalidArgument (a ValueError)
    std::out_of_range     -> OutOfRange      (an IndexError)
    std::runtime_error    -> BackendRuntimeError (a RuntimeError)
    std::logic_error      -> LogicError

There is no CPU fallback: if libpvhip.so is missing, importing the library raises.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libpvhip.so"
HEADER_PATH = Path(__file__).resolve().parent.parent / "include" / "pv_hip.h"

GEOMETRIES = {"infinite": 0, "periodic": 1, "disk": 2}
PRECISIONS = {"fp64": 0, "mixed": 1}


class InvalidArgument(ValueError):
    """std::invalid_argument."""


class OutOfRange(IndexError):
    """std::out_of_range."""


class BackendRuntimeError(RuntimeError):
    """std::runtime_error."""


class LogicError(Exception):
    """std::logic_error."""


_ERRORS = {1: InvalidArgument, 2: OutOfRange, 3: BackendRuntimeError, 4: LogicError}


class PvConfig(ctypes.Structure):
    _fields_ = [
        ("geometry", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("core_radius", ctypes.c_double),
        ("box_length_x", ctypes.c_double),
        ("box_length_y", ctypes.c_double),
        ("image_layers", ctypes.c_int),
        ("disk_radius", ctypes.c_double),
        ("device", ctypes.c_int),
    ]


class PvStepResult(ctypes.Structure):
    _fields_ = [
        ("accepted_time_step", ctypes.c_double),
        ("suggested_time_step", ctypes.c_double),
        ("normalized_error", ctypes.c_double),
        ("rejected_steps", ctypes.c_uint),
    ]


class PvInvariants(ctypes.Structure):
    _fields_ = [
        ("circulation", ctypes.c_double),
        ("linear_impulse_x", ctypes.c_double),
        ("linear_impulse_y", ctypes.c_double),
        ("angular_impulse", ctypes.c_double),
        ("hamiltonian", ctypes.c_double),
    ]


class PvInstanceConfig(ctypes.Structure):
    _fields_ = [
        ("geometry", ctypes.c_int),
        ("image_layers", ctypes.c_int),
        ("core_radius", ctypes.c_double),
        ("box_length", ctypes.c_double),
        ("disk_radius", ctypes.c_double),
    ]


class PvLaunchRule(ctypes.Structure):
    _fields_ = [
        ("geometry", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("n_max", ctypes.c_size_t),
        ("targets_per_thread", ctypes.c_int),
        ("source_splits", ctypes.c_int),
        ("tile", ctypes.c_int),
        ("symmetric", ctypes.c_int),
    ]


class PvBatchRule(ctypes.Structure):
    _fields_ = [
        ("geometry", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("n_max", ctypes.c_size_t),
        ("targets_per_lane", ctypes.c_int),
        ("chunk_items", ctypes.c_int),
    ]


class PvSlotState(ctypes.Structure):
    _fields_ = [
        ("slot", ctypes.c_int),
        ("state", ctypes.c_int),
        ("t", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("last_error", ctypes.c_double),
        ("accepted", ctypes.c_uint),
        ("rejected", ctypes.c_uint),
        ("failure", ctypes.c_uint),
        ("n", ctypes.c_size_t),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_szp = ctypes.POINTER(ctypes.c_size_t)
_vp = ctypes.c_void_p

# name -> (restype, argtypes)
_SIGNATURES = {
    "pv_last_error": (ctypes.c_char_p, []),
    "pv_create": (ctypes.c_int, [ctypes.POINTER(PvConfig), ctypes.POINTER(_vp)]),
    "pv_destroy": (None, [_vp]),
    "pv_set_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pv_evaluate_range": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_size_t, _dp, _dp,
                                         ctypes.c_size_t, ctypes.c_size_t]),
    "pv_evaluate": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_size_t, _dp, _dp]),
    "pv_hamiltonian": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_size_t, _dp]),
    "pv_invariants_of": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_size_t,
                                        ctypes.POINTER(PvInvariants)]),
    "pv_upload": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_size_t]),
    "pv_download": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_size_t]),
    "pv_state_size": (ctypes.c_int, [_vp, _szp]),
    "pv_evaluate_device_state": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_size_t]),
    "pv_rk4_step": (ctypes.c_int, [_vp, ctypes.c_double]),
    "pv_rk4_advance": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int]),
    "pv_dopri5_step": (ctypes.c_int, [_vp] + [ctypes.c_double] * 5 +
                       [ctypes.POINTER(PvStepResult)]),
    "pv_invalidate_fsal": (None, [_vp]),
    "pv_state_invariants": (ctypes.c_int, [_vp, ctypes.POINTER(PvInvariants)]),
    "pv_dipole_candidates": (ctypes.c_int, [_vp, ctypes.c_double, _u64p, _u64p, _dp,
                                            ctypes.c_size_t, _szp]),
    "pv_runtime_details": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
    "pv_evaluate_resident": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pv_synchronize": (ctypes.c_int, [_vp]),
    "pv_set_launch": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "pv_set_launch_shape": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int]),
    "pv_set_launch_table": (ctypes.c_int, [_vp, ctypes.POINTER(PvLaunchRule), ctypes.c_size_t]),
    "pv_resolve_launch": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_size_t,
                                         ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]),
    "pv_set_persistent": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
    "pv_persistent_blocks": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)]),
    "pv_profile_begin": (ctypes.c_int, [_vp]),
    "pv_profile_end": (ctypes.c_int, [_vp, _dp, ctypes.POINTER(ctypes.c_int)]),
    "pv_calibration_copy": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, _dp]),
    "pv_evaluate_device_noalloc": (ctypes.c_int, [ctypes.POINTER(PvConfig), _vp, _vp, _vp,
                                                  ctypes.c_size_t, _vp, _vp, _vp]),
    "pv_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                                       ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "pv_batch_destroy": (None, [_vp]),
    "pv_batch_set_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pv_batch_set_symmetric": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pv_batch_set_byte_limit": (ctypes.c_int, [_vp, ctypes.c_size_t]),
    "pv_batch_device_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    "pv_dipole_process": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_size_t)]),
    "pv_dipole_set_random": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_char_p]),
    "pv_dipole_random_state": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t,
                                              ctypes.POINTER(ctypes.c_size_t)]),
    "pv_batch_symmetric": (ctypes.c_int, [_vp]),
    "pv_batch_set_launch_table": (ctypes.c_int, [_vp, ctypes.POINTER(PvBatchRule),
                                                 ctypes.c_size_t]),
    "pv_batch_resolve_launch": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                                               ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_int)]),
    "pv_batch_instance_bytes": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                                               ctypes.c_int, _szp]),
    "pv_batch_bytes_per_vortex": (ctypes.c_size_t, []),
    "pv_batch_load": (ctypes.c_int, [_vp, ctypes.POINTER(PvInstanceConfig), _szp,
                                     ctypes.c_size_t, _dp, _dp, _dp]),
    "pv_batch_rk4": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    "pv_batch_read": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_size_t]),
    "pv_batch_evaluate": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_size_t]),
    "pv_batch_failures": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint),
                                         ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_size_t]),
    "pv_batch_admit": (ctypes.c_int, [_vp, ctypes.POINTER(PvInstanceConfig), _dp, _dp, _dp,
                                      ctypes.c_size_t] + [ctypes.c_double] * 6 +
                       [ctypes.POINTER(ctypes.c_int)]),
    "pv_batch_attempt": (ctypes.c_int, [_vp]),
    "pv_batch_attempts": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pv_batch_poll": (ctypes.c_int, [_vp, ctypes.POINTER(PvSlotState), ctypes.c_size_t, _szp]),
    "pv_batch_fetch": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_size_t]),
    "pv_batch_release": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pv_batch_failure_message": (ctypes.c_char_p, [ctypes.c_uint]),
}

_lib = None


def load_library(path: Path = LIB_PATH):
    """Loads libpvhip.so (raises if it has not been built: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not Path(path).exists():
        raise RuntimeError(f"{path} is missing; build it with `python -m pointvortex_amd.build`")
    lib = ctypes.CDLL(str(path))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(status):
    if status:
        msg = _lib.pv_last_error().decode()
        raise _ERRORS.get(status, BackendRuntimeError)(msg)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _ptr(a):
    return a.ctypes.data_as(_dp)


@dataclass
class SimParams:
    """The geometry/integrator fields of SimParams (src/params.h:37-78) this backend reads."""

    boundaryCondition: str = "infinite"
    coreRadius: float = 0.0
    boxLengthX: float = 2.0
    boxLengthY: float = 2.0
    diskRadius: float = 1.0
    periodicImageLayers: int = 8
    absoluteTolerance: float = 1.0e-10
    relativeTolerance: float = 1.0e-8
    minimumTimeStep: float = 1.0e-12
    maximumTimeStep: float = 0.1
    precision: str = "fp64"
    device: int = 0


@dataclass
class VortexSystem:
    """src/vortex.h:15-38: structure-of-arrays FP64 state.  Like the reference's struct of
    std::vectors it owns its arrays: construction copies them, so downloadDeviceState into a
    system never writes through to the arrays it was built from."""

    x: np.ndarray = field(default_factory=lambda: np.zeros(0))
    y: np.ndarray = field(default_factory=lambda: np.zeros(0))
    circulation: np.ndarray = field(default_factory=lambda: np.zeros(0))

    def __post_init__(self):
        self.x = np.array(self.x, dtype=np.float64, order="C", copy=True)
        self.y = np.array(self.y, dtype=np.float64, order="C", copy=True)
        self.circulation = np.array(self.circulation, dtype=np.float64, order="C", copy=True)

    def size(self):
        return len(self.x)

    def copy(self):
        return VortexSystem(self.x.copy(), self.y.copy(), self.circulation.copy())


@dataclass
class VectorField:
    """src/vortex.h:40-51 (velocities and stage fields)."""

    x: np.ndarray
    y: np.ndarray


@dataclass
class StepResult:
    """src/compute.h:11-16."""

    acceptedTimeStep: float
    suggestedTimeStep: float
    normalizedError: float
    rejectedSteps: int


@dataclass
class Invariants:
    """src/compute.h:78-81."""

    circulation: float
    linearImpulseX: float
    linearImpulseY: float
    angularImpulse: float
    hamiltonian: float


class HipKernel:
    """The VelocityKernel of point_vortex_hip (src/compute.h:18-36), device stepping included."""

    def __init__(self, params: Optional[SimParams] = None, **overrides):
        lib = load_library()
        p = params or SimParams()
        for k, v in overrides.items():
            setattr(p, k, v)
        if p.boundaryCondition not in GEOMETRIES:
            raise InvalidArgument("unsupported boundary condition: " + p.boundaryCondition)
        self.params = p
        cfg = PvConfig(GEOMETRIES[p.boundaryCondition], PRECISIONS[p.precision],
                       float(p.coreRadius), float(p.boxLengthX), float(p.boxLengthY),
                       int(p.periodicImageLayers), float(p.diskRadius), int(p.device))
        h = _vp()
        _check(lib.pv_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.pv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration (no reference counterpart)
    def set_precision(self, precision: str):
        _check(_lib.pv_set_precision(self._h, PRECISIONS[precision]))
        self.params.precision = precision

    def set_launch(self, targets_per_thread=0, source_splits=0, tile=0, symmetric=-1):
        """Explicit launch shape of the pair kernels (0 = automatic / default; symmetric -1 =
        not forced, 0 = split path, 1 = symmetric path for whole-population evaluations)."""
        _check(_lib.pv_set_launch_shape(self._h, int(targets_per_thread), int(source_splits),
                                        int(tile), int(symmetric)))

    def set_persistent(self, mode: int = -1, targets_per_block: int = 0,
                      max_wait_seconds: float = 0.0):
        """RK4 steps in one cooperative launch for small populations: mode -1 automatic
        (N <= 4096 unless a launch shape was forced), 1 whenever it fits, 0 never;
        targets_per_block 8, 16 or 32 (0 = 16)."""
        _check(_lib.pv_set_persistent(self._h, int(mode), int(targets_per_block),
                                      float(max_wait_seconds)))

    def persistent_blocks(self, n: int) -> int:
        """Blocks of the persistent RK4 launch for n vortices (0: per-stage launches)."""
        b = ctypes.c_int()
        _check(_lib.pv_persistent_blocks(self._h, int(n), ctypes.byref(b)))
        return b.value

    def set_launch_table(self, rules):
        """rules: iterable of (geometry, precision, n_max, targets_per_thread, source_splits,
        tile[, symmetric]); the first rule matching the handle's geometry, precision and N
        applies."""
        rules = [tuple(r) + (0,) * (7 - len(r)) for r in rules]
        arr = (PvLaunchRule * max(1, len(rules)))(
            *[PvLaunchRule(GEOMETRIES[g], PRECISIONS[p], int(n), int(t), int(s), int(tl),
                           int(sy)) for g, p, n, t, s, tl, sy in rules])
        _check(_lib.pv_set_launch_table(self._h, arr, len(rules)))

    def resolved_launch(self, targets, n):
        """(targets_per_thread, source_splits, tile, symmetric) an evaluation would use."""
        t, s, tl, sy = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(_lib.pv_resolve_launch(self._h, int(targets), int(n), ctypes.byref(t),
                                      ctypes.byref(s), ctypes.byref(tl), ctypes.byref(sy)))
        return t.value, s.value, tl.value, sy.value

    # -- pure evaluation (src/compute.h:21-27)
    def evaluateRange(self, x, y, circulation, begin, end, velocity: Optional[VectorField] = None):
        x, y, g = _f64(x), _f64(y), _f64(circulation)
        n = len(x)
        if len(y) != n or len(g) != n:
            raise InvalidArgument("vortex arrays have different lengths")
        if velocity is None or len(velocity.x) != n:
            velocity = VectorField(np.zeros(n), np.zeros(n))
        _check(_lib.pv_evaluate_range(self._h, _ptr(x), _ptr(y), _ptr(g), n, _ptr(velocity.x),
                                      _ptr(velocity.y), int(begin), int(end)))
        return velocity

    def evaluate(self, vortices: VortexSystem) -> VectorField:
        x, y, g = vortices.x, vortices.y, vortices.circulation
        n = len(x)
        if len(y) != n or len(g) != n:
            raise InvalidArgument("vortex arrays have different lengths")
        out = VectorField(np.zeros(n), np.zeros(n))
        _check(_lib.pv_evaluate(self._h, _ptr(_f64(x)), _ptr(_f64(y)), _ptr(_f64(g)), n,
                                _ptr(out.x), _ptr(out.y)))
        return out

    def hamiltonian(self, vortices: VortexSystem) -> float:
        out = ctypes.c_double(0.0)
        _check(_lib.pv_hamiltonian(self._h, _ptr(vortices.x), _ptr(vortices.y),
                                   _ptr(vortices.circulation), vortices.size(), ctypes.byref(out)))
        return out.value

    def invariantsOf(self, vortices: VortexSystem) -> Invariants:
        out = PvInvariants()
        _check(_lib.pv_invariants_of(self._h, _ptr(vortices.x), _ptr(vortices.y),
                                     _ptr(vortices.circulation), vortices.size(),
                                     ctypes.byref(out)))
        return _invariants(out)

    # -- device stepping (src/compute.h:29-35)
    @staticmethod
    def supportsDeviceStepping():
        return True

    def uploadDeviceState(self, vortices: VortexSystem):
        _check(_lib.pv_upload(self._h, _ptr(vortices.x), _ptr(vortices.y),
                              _ptr(vortices.circulation), vortices.size()))

    def downloadDeviceState(self, vortices: VortexSystem):
        n = vortices.size()
        _check(_lib.pv_download(self._h, _ptr(vortices.x), _ptr(vortices.y), None, n))

    def downloadDeviceSystem(self) -> VortexSystem:
        """The whole resident system, circulations included (after processDipoles they are
        no longer the uploaded ones)."""
        n = self.stateSize()
        s = VortexSystem(np.zeros(n), np.zeros(n), np.zeros(n))
        if n:
            _check(_lib.pv_download(self._h, _ptr(s.x), _ptr(s.y), _ptr(s.circulation), n))
        return s

    def stateSize(self) -> int:
        n = ctypes.c_size_t(0)
        _check(_lib.pv_state_size(self._h, ctypes.byref(n)))
        return n.value

    def evaluateDeviceState(self) -> VectorField:
        n = self.stateSize()
        out = VectorField(np.zeros(n), np.zeros(n))
        _check(_lib.pv_evaluate_device_state(self._h, _ptr(out.x), _ptr(out.y), n))
        return out

    def deviceRk4Step(self, dt: float):
        _check(_lib.pv_rk4_step(self._h, float(dt)))

    def deviceRk4Advance(self, dt: float, steps: int):
        _check(_lib.pv_rk4_advance(self._h, float(dt), int(steps)))

    def deviceDopri5Step(self, dt, atol, rtol, min_dt, max_dt) -> StepResult:
        r = PvStepResult()
        _check(_lib.pv_dopri5_step(self._h, float(dt), float(atol), float(rtol), float(min_dt),
                                   float(max_dt), ctypes.byref(r)))
        return StepResult(r.accepted_time_step, r.suggested_time_step, r.normalized_error,
                          int(r.rejected_steps))

    def invalidateDeviceDerivative(self):
        _lib.pv_invalidate_fsal(self._h)

    def deviceInvariants(self) -> Invariants:
        out = PvInvariants()
        _check(_lib.pv_state_invariants(self._h, ctypes.byref(out)))
        return _invariants(out)

    def dipoleCandidates(self, distance: float, cap: Optional[int] = None):
        """Unordered (first, second, d^2) of src/dipole.cpp:87-119 on the device state."""
        n = self.stateSize()
        cap = max(16, n) if cap is None else cap
        first = np.zeros(cap, dtype=np.uint64)
        second = np.zeros(cap, dtype=np.uint64)
        d2 = np.zeros(cap)
        count = ctypes.c_size_t(0)
        _check(_lib.pv_dipole_candidates(self._h, float(distance), first.ctypes.data_as(_u64p),
                                         second.ctypes.data_as(_u64p), _ptr(d2), cap,
                                         ctypes.byref(count)))
        if count.value > cap:
            return self.dipoleCandidates(distance, count.value)
        c = count.value
        return first[:c].copy(), second[:c].copy(), d2[:c].copy()

    REINJECTION = {"none": 0, "independent": 1, "paired": 2}

    def processDipoles(self, distance: float, reinjection: str = "none") -> int:
        """DipoleManager::process (src/dipole.cpp:42-125) on the device state: the search on
        the device, then the reference's selection, compaction and seeded reinjection; the
        resident state may change size.  Returns the number of removed pairs."""
        events = ctypes.c_size_t(0)
        _check(_lib.pv_dipole_process(self._h, float(distance), self.REINJECTION[reinjection],
                                      ctypes.byref(events)))
        return events.value

    def setDipoleRandom(self, seed: int, state: str = ""):
        """The reinjection generator: seeded like DipoleManager (seed ^ 0xd1b54a32d192ed03),
        or restored from a checkpoint's generator text."""
        _check(_lib.pv_dipole_set_random(self._h, int(seed) & (2 ** 64 - 1), state.encode()))

    def dipoleRandomState(self) -> str:
        need = ctypes.c_size_t(0)
        _check(_lib.pv_dipole_random_state(self._h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _check(_lib.pv_dipole_random_state(self._h, buf, need.value, ctypes.byref(need)))
        return buf.value.decode()

    # -- benchmark helpers (device-resident, no host round trip)
    def evaluateResident(self, repeats=1):
        _check(_lib.pv_evaluate_resident(self._h, int(repeats)))

    def synchronize(self):
        _check(_lib.pv_synchronize(self._h))

    def profileBegin(self):
        _check(_lib.pv_profile_begin(self._h))

    def profileEnd(self):
        """(summed velocity-kernel milliseconds, launches) since profileBegin."""
        ms = ctypes.c_double(0.0)
        cnt = ctypes.c_int(0)
        _check(_lib.pv_profile_end(self._h, ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value


def _invariants(c: PvInvariants) -> Invariants:
    return Invariants(c.circulation, c.linear_impulse_x, c.linear_impulse_y, c.angular_impulse,
                      c.hamiltonian)


def makeBackendKernel(params: SimParams) -> HipKernel:
    """src/backend.h:16."""
    return HipKernel(params)


def computeInvariants(vortices: VortexSystem, kernel: HipKernel) -> Invariants:
    """src/compute.cpp:157-163, moments and Hamiltonian on the device."""
    return kernel.invariantsOf(vortices)


def backendName() -> str:
    """src/backend.h:13."""
    return "HIP"


def backendRuntimeDetails(device: int = 0) -> str:
    """src/backend.h:14."""
    lib = load_library()
    buf = ctypes.create_string_buffer(1024)
    _check(lib.pv_runtime_details(int(device), buf, len(buf)))
    return buf.value.decode()


NOALLOC_SCRATCH_BYTES = 16 << 20
NOALLOC_MAX_VORTICES = 1 << 22


def evaluate_device_noalloc(params: SimParams, x_ptr: int, y_ptr: int, g_ptr: int, n: int,
                            u_ptr: int, v_ptr: int, stream: int = 0):
    """pv_evaluate_device_noalloc: VelocityKernel::evaluate (src/compute.cpp:60-69) of a
    population already in device memory (x, y, circulation, u, v are device addresses), with
    no device allocation: the kernels' scratch is a fixed per-GPU array of the library."""
    lib = load_library()
    p = params
    cfg = PvConfig(GEOMETRIES[p.boundaryCondition], PRECISIONS[p.precision],
                   float(p.coreRadius), float(p.boxLengthX), float(p.boxLengthY),
                   int(p.periodicImageLayers), float(p.diskRadius), int(p.device))
    _check(lib.pv_evaluate_device_noalloc(ctypes.byref(cfg), x_ptr, y_ptr, g_ptr, int(n),
                                          u_ptr, v_ptr, stream or None))


def calibration_copy(nbytes: int, repeats: int = 10, device: int = 0) -> float:
    """Average milliseconds of a plain device copy kernel of nbytes (counter calibration)."""
    lib = load_library()
    ms = ctypes.c_double(0.0)
    _check(lib.pv_calibration_copy(int(device), int(nbytes), int(repeats), ctypes.byref(ms)))
    return ms.value


def declared_symbols(header: Path = HEADER_PATH):
    """Function names declared in include/pv_hip.h."""
    import re

    text = header.read_text()
    return sorted(set(re.findall(r"^\s*(?:pv_status|void|const char \*|size_t|int)\s*(pv_\w+)\s*\(",
                                 text, re.M)))
