"""ctypes bindings for the CPU checkers (test infrastructure only).

* ``Oracle`` wraps ``oracle/_build/libpvoracle.so``: the repo's own C FP64 restatement of
  the reference's velocity backend (``oracle/pv_oracle.c``).
* ``Reference`` wraps ``oracle/_ref/libpvref.so``: the reference's own CPU kernels compiled
  from its sources, behind ``oracle/ref_shim.cpp``.  Present only where it was built.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this
module.  The product (``pointvortex_amd``) never does.
"""
import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "_build" / "libpvoracle.so"
REF_LIB = HERE / "_ref" / "libpvref.so"
REF_CPU_EXE = HERE / "_ref" / "point_vortex_cpu"
REF_HIP_EXE = HERE / "_ref" / "point_vortex_hip"

GEOMETRIES = {"infinite": 0, "periodic": 1, "disk": 2}
ERRORS = {1: ValueError, 2: IndexError, 3: RuntimeError, 4: AssertionError}

_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_szp = ctypes.POINTER(ctypes.c_size_t)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class KernelSpec(ctypes.Structure):
    """Mirror of ``pvo_kernel`` in pv_oracle.c."""

    _fields_ = [
        ("geometry", ctypes.c_int),
        ("image_layers", ctypes.c_int),
        ("core_radius", ctypes.c_double),
        ("length_x", ctypes.c_double),
        ("length_y", ctypes.c_double),
        ("radius", ctypes.c_double),
    ]


def spec(geometry="infinite", core_radius=0.0, box_length=2.0, image_layers=8, radius=1.0,
         box_length_y=None):
    return KernelSpec(GEOMETRIES[geometry], int(image_layers), float(core_radius),
                      float(box_length), float(box_length if box_length_y is None else box_length_y),
                      float(radius))


class Oracle:
    """The repo's own CPU FP64 restatement (pv_oracle.c)."""

    def __init__(self, path=ORACLE_LIB, threads=None):
        if threads is not None:
            os.environ["OMP_NUM_THREADS"] = str(threads)
        self.lib = ctypes.CDLL(str(path))
        L = self.lib
        L.pvo_last_error.restype = ctypes.c_char_p
        ks = ctypes.POINTER(KernelSpec)
        L.pvo_evaluate_range.argtypes = [ks, _dp, _dp, _dp, ctypes.c_size_t, _dp, _dp,
                                         ctypes.c_size_t, ctypes.c_size_t]
        L.pvo_evaluate.argtypes = [ks, _dp, _dp, _dp, ctypes.c_size_t, _dp, _dp]
        L.pvo_hamiltonian.argtypes = [ks, _dp, _dp, _dp, ctypes.c_size_t, _dp]
        L.pvo_invariants.argtypes = [ks, _dp, _dp, _dp, ctypes.c_size_t, _dp]
        L.pvo_moments.argtypes = [_dp, _dp, _dp, ctypes.c_size_t, _dp]
        L.pvo_check_kernel.argtypes = [ks]
        L.pvo_integrator_create.restype = ctypes.c_void_p
        L.pvo_integrator_create.argtypes = [ctypes.c_size_t]
        L.pvo_integrator_destroy.argtypes = [ctypes.c_void_p]
        L.pvo_integrator_invalidate.argtypes = [ctypes.c_void_p]
        L.pvo_rk4_step.argtypes = [ctypes.c_void_p, ks, _dp, _dp, _dp, ctypes.c_size_t,
                                   ctypes.c_double]
        L.pvo_dopri5_step.argtypes = [ctypes.c_void_p, ks, _dp, _dp, _dp, ctypes.c_size_t,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, _dp]
        L.pvo_dipole_candidates.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, _dp, _dp, _dp,
                                            ctypes.c_size_t, _u64p, _u64p, _dp, ctypes.c_size_t,
                                            _szp]

    def _check(self, status):
        if status:
            raise ERRORS.get(status, RuntimeError)(self.lib.pvo_last_error().decode())

    def check_kernel(self, k):
        self._check(self.lib.pvo_check_kernel(ctypes.byref(k)))

    def evaluate(self, k, x, y, g):
        x, y, g = _f64(x), _f64(y), _f64(g)
        u = np.zeros_like(x)
        v = np.zeros_like(x)
        self._check(self.lib.pvo_evaluate(ctypes.byref(k), _ptr(x), _ptr(y), _ptr(g), len(x),
                                          _ptr(u), _ptr(v)))
        return u, v

    def evaluate_range(self, k, x, y, g, begin, end):
        x, y, g = _f64(x), _f64(y), _f64(g)
        u = np.zeros_like(x)
        v = np.zeros_like(x)
        self._check(self.lib.pvo_evaluate_range(ctypes.byref(k), _ptr(x), _ptr(y), _ptr(g),
                                                len(x), _ptr(u), _ptr(v), begin, end))
        return u, v

    def hamiltonian(self, k, x, y, g):
        x, y, g = _f64(x), _f64(y), _f64(g)
        out = np.zeros(1)
        self._check(self.lib.pvo_hamiltonian(ctypes.byref(k), _ptr(x), _ptr(y), _ptr(g), len(x),
                                             _ptr(out)))
        return float(out[0])

    def invariants(self, k, x, y, g):
        x, y, g = _f64(x), _f64(y), _f64(g)
        out = np.zeros(5)
        self._check(self.lib.pvo_invariants(ctypes.byref(k), _ptr(x), _ptr(y), _ptr(g), len(x),
                                            _ptr(out)))
        return out

    def integrator(self, n):
        return OracleIntegrator(self, n)

    def dipole_candidates(self, geometry, box_length, radius, distance, x, y, g, cap=None):
        x, y, g = _f64(x), _f64(y), _f64(g)
        n = len(x)
        cap = cap if cap is not None else max(16, n)
        first = np.zeros(cap, dtype=np.uint64)
        second = np.zeros(cap, dtype=np.uint64)
        d2 = np.zeros(cap)
        count = ctypes.c_size_t(0)
        self._check(self.lib.pvo_dipole_candidates(
            GEOMETRIES[geometry], box_length, box_length, radius, distance, _ptr(x), _ptr(y),
            _ptr(g), n, first.ctypes.data_as(_u64p), second.ctypes.data_as(_u64p), _ptr(d2), cap,
            ctypes.byref(count)))
        c = count.value
        if c > cap:
            return self.dipole_candidates(geometry, box_length, radius, distance, x, y, g, cap=c)
        return first[:c].copy(), second[:c].copy(), d2[:c].copy()


class OracleIntegrator:
    """RungeKuttaIntegrator restatement (src/timestep.h:18-34)."""

    def __init__(self, oracle, n):
        self.o = oracle
        self.h = oracle.lib.pvo_integrator_create(n)

    def __del__(self):
        if getattr(self, "h", None):
            self.o.lib.pvo_integrator_destroy(self.h)
            self.h = None

    def invalidate(self):
        self.o.lib.pvo_integrator_invalidate(self.h)

    def rk4_step(self, k, x, y, g, dt):
        """Advances x, y (float64 arrays) in place."""
        self.o._check(self.o.lib.pvo_rk4_step(self.h, ctypes.byref(k), _ptr(x), _ptr(y), _ptr(g),
                                              len(x), dt))

    def dopri5_step(self, k, x, y, g, dt, atol, rtol, min_dt, max_dt):
        out = np.zeros(4)
        self.o._check(self.o.lib.pvo_dopri5_step(self.h, ctypes.byref(k), _ptr(x), _ptr(y),
                                                 _ptr(g), len(x), dt, atol, rtol, min_dt, max_dt,
                                                 _ptr(out)))
        return out[0], out[1], out[2], int(out[3])


class Reference:
    """The reference's own CPU kernels (oracle/_ref/libpvref.so)."""

    def __init__(self, path=REF_LIB):
        self.lib = ctypes.CDLL(str(path))
        L = self.lib
        L.pvr_last_error.restype = ctypes.c_char_p
        geo = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
               ctypes.c_double]
        L.pvr_evaluate.argtypes = geo + [_dp, _dp, _dp, ctypes.c_size_t, _dp, _dp]
        L.pvr_evaluate_range.argtypes = geo + [_dp, _dp, _dp, ctypes.c_size_t, _dp, _dp,
                                               ctypes.c_size_t, ctypes.c_size_t]
        L.pvr_invariants.argtypes = geo + [_dp, _dp, _dp, ctypes.c_size_t, _dp]
        L.pvr_rk4.argtypes = geo + [_dp, _dp, _dp, ctypes.c_size_t, ctypes.c_double, ctypes.c_int]
        L.pvr_dopri5_run.argtypes = geo + [_dp, _dp, _dp, ctypes.c_size_t] + [ctypes.c_double] * 6 \
            + [_dp, ctypes.c_size_t, _szp]
        L.pvr_dipole_process.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_uint64, _dp, _dp, _dp, ctypes.c_size_t, _szp,
                                         _szp]

    @staticmethod
    def available():
        return REF_LIB.exists()

    def _check(self, status):
        if status:
            raise ERRORS.get(status, RuntimeError)(self.lib.pvr_last_error().decode())

    @staticmethod
    def _geo(k):
        return (k.geometry, k.core_radius, k.length_x, k.length_y, k.image_layers, k.radius)

    def evaluate(self, k, x, y, g):
        x, y, g = _f64(x), _f64(y), _f64(g)
        u = np.zeros_like(x)
        v = np.zeros_like(x)
        self._check(self.lib.pvr_evaluate(*self._geo(k), _ptr(x), _ptr(y), _ptr(g), len(x),
                                          _ptr(u), _ptr(v)))
        return u, v

    def evaluate_range(self, k, x, y, g, begin, end):
        x, y, g = _f64(x), _f64(y), _f64(g)
        u = np.zeros_like(x)
        v = np.zeros_like(x)
        self._check(self.lib.pvr_evaluate_range(*self._geo(k), _ptr(x), _ptr(y), _ptr(g), len(x),
                                                _ptr(u), _ptr(v), begin, end))
        return u, v

    def invariants(self, k, x, y, g):
        x, y, g = _f64(x), _f64(y), _f64(g)
        out = np.zeros(5)
        self._check(self.lib.pvr_invariants(*self._geo(k), _ptr(x), _ptr(y), _ptr(g), len(x),
                                            _ptr(out)))
        return out

    def rk4(self, k, x, y, g, dt, steps):
        x, y, g = _f64(x).copy(), _f64(y).copy(), _f64(g)
        self._check(self.lib.pvr_rk4(*self._geo(k), _ptr(x), _ptr(y), _ptr(g), len(x), dt, steps))
        return x, y

    def dopri5_run(self, k, x, y, g, dt, end, atol, rtol, min_dt, max_dt, cap=100000):
        x, y, g = _f64(x).copy(), _f64(y).copy(), _f64(g)
        rec = np.zeros(4 * cap)
        taken = ctypes.c_size_t(0)
        self._check(self.lib.pvr_dopri5_run(*self._geo(k), _ptr(x), _ptr(y), _ptr(g), len(x), dt,
                                            end, atol, rtol, min_dt, max_dt, _ptr(rec), cap,
                                            ctypes.byref(taken)))
        return x, y, rec[:4 * min(cap, taken.value)].reshape(-1, 4)

    def dipole_process(self, geometry, box_length, radius, distance, reinjection, seed, x, y, g):
        x, y, g = _f64(x).copy(), _f64(y).copy(), _f64(g).copy()
        n_new = ctypes.c_size_t(0)
        events = ctypes.c_size_t(0)
        self._check(self.lib.pvr_dipole_process(GEOMETRIES[geometry], box_length, box_length,
                                                radius, distance, reinjection, seed, _ptr(x),
                                                _ptr(y), _ptr(g), len(x), ctypes.byref(n_new),
                                                ctypes.byref(events)))
        m = n_new.value
        return x[:m], y[:m], g[:m], events.value
