"""Ensembles of independent vortex systems on one GPU (the pv_batch_* C-ABI).

Two uses:
* fixed-step ensembles -- many instances of uneven N and any geometry advanced together with
  RK4 (one persistent launch per stage, steps replayed from a captured hipGraph);
* adaptive ensembles with admission while others run -- each run has its own DOPRI5 state
  and step-size controller on the device; runs are admitted as vortex ranges and instance
  slots free up, and released when they finish (see pointvortex_amd.serve).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from .backend import (GEOMETRIES, PRECISIONS, PvBatchRule, PvInstanceConfig, PvSlotState,
                      _check, _f64, _ptr, _szp, _vp, load_library)

STATES = {0: "running", 1: "finished", 2: "failed", 3: "idle"}


@dataclass
class InstanceConfig:
    """Geometry of one instance: the SimParams fields a kernel reads (src/params.h:44-57)."""

    geometry: str = "infinite"
    core_radius: float = 0.0
    box_length: float = 2.0
    disk_radius: float = 1.0
    image_layers: int = 8

    def c(self):
        return PvInstanceConfig(GEOMETRIES[self.geometry], int(self.image_layers),
                                float(self.core_radius), float(self.box_length),
                                float(self.disk_radius))


@dataclass
class SlotState:
    slot: int
    state: str
    t: float
    dt: float
    last_error: float
    accepted: int
    rejected: int
    failure: str
    n: int


def bytes_per_vortex() -> int:
    return int(load_library().pv_batch_bytes_per_vortex())


class Ensemble:
    def __init__(self, capacity: int, max_instances: int, precision: str = "fp64",
                 device: int = 0):
        self.lib = load_library()
        h = _vp()
        _check(self.lib.pv_batch_create(int(device), PRECISIONS[precision], int(capacity),
                                        int(max_instances), ctypes.byref(h)))
        self._h = h
        self.capacity = capacity
        self.max_instances = max_instances
        self.counts: List[int] = []

    def close(self):
        if getattr(self, "_h", None):
            self.lib.pv_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, precision: str):
        """Fixed-step ensembles: every instance.  Admitted runs: the runs admitted next (each
        run keeps its own; FP64 and mixed runs advance in the same launches)."""
        _check(self.lib.pv_batch_set_precision(self._h, PRECISIONS[precision]))

    def set_byte_limit(self, nbytes: int):
        """Admit runs only while everything the ensemble holds on the device stays within
        `nbytes` (0: no limit)."""
        _check(self.lib.pv_batch_set_byte_limit(self._h, int(nbytes)))

    def device_bytes(self) -> int:
        out = ctypes.c_size_t(0)
        _check(self.lib.pv_batch_device_bytes(self._h, ctypes.byref(out)))
        return out.value

    def set_symmetric(self, mode: int):
        """1: each unordered pair once over tile pairs; 0: ordered 64-target items; -1
        (default): symmetric for fixed-step ensembles, ordered for admitted runs."""
        _check(self.lib.pv_batch_set_symmetric(self._h, int(mode)))

    @property
    def symmetric(self) -> bool:
        return bool(self.lib.pv_batch_symmetric(self._h))

    def set_launch_table(self, rules):
        """rules: iterable of (geometry, precision, n_max, targets_per_lane, chunk_items) for
        the symmetric decomposition (tiles of 64 targets_per_lane vortices); [] restores the
        defaults (2 targets per lane, chunks of 4)."""
        rules = list(rules)
        arr = (PvBatchRule * max(1, len(rules)))(*[
            PvBatchRule(GEOMETRIES[g], PRECISIONS[p], int(n), int(t), int(c))
            for g, p, n, t, c in rules])
        _check(self.lib.pv_batch_set_launch_table(self._h, arr, len(rules)))

    def resolve_launch(self, geometry: str, precision: str, n: int):
        """(targets_per_lane, chunk_items) an instance of n vortices takes."""
        t, c = ctypes.c_int(0), ctypes.c_int(0)
        _check(self.lib.pv_batch_resolve_launch(self._h, GEOMETRIES[geometry],
                                                PRECISIONS[precision], int(n), ctypes.byref(t),
                                                ctypes.byref(c)))
        return t.value, c.value

    def instance_bytes(self, geometry: str, precision: str, n: int, symmetric: bool = True):
        """Device bytes one admitted run of n vortices adds beyond its per-vortex arrays (work
        items, and in the symmetric decomposition its partial sums at its tile)."""
        out = ctypes.c_size_t(0)
        _check(self.lib.pv_batch_instance_bytes(self._h, GEOMETRIES[geometry],
                                                PRECISIONS[precision], int(n),
                                                int(bool(symmetric)), ctypes.byref(out)))
        return out.value

    # -- fixed-step ensembles
    def load(self, configs: Sequence[InstanceConfig], xs, ys, gs):
        counts = [len(x) for x in xs]
        arr = (PvInstanceConfig * max(1, len(configs)))(*[c.c() for c in configs])
        cnt = (ctypes.c_size_t * max(1, len(counts)))(*counts)
        x = _f64(np.concatenate(xs)) if counts else np.zeros(0)
        y = _f64(np.concatenate(ys)) if counts else np.zeros(0)
        g = _f64(np.concatenate(gs)) if counts else np.zeros(0)
        _check(self.lib.pv_batch_load(self._h, arr, cnt, len(configs), _ptr(x), _ptr(y),
                                      _ptr(g)))
        self.counts = counts

    def rk4(self, dt: float, steps: int, use_graph: bool = True):
        _check(self.lib.pv_batch_rk4(self._h, float(dt), int(steps), int(bool(use_graph))))

    def _split(self, a):
        out, o = [], 0
        for c in self.counts:
            out.append(a[o:o + c].copy())
            o += c
        return out

    def read(self):
        total = sum(self.counts)
        x, y = np.zeros(total), np.zeros(total)
        _check(self.lib.pv_batch_read(self._h, _ptr(x), _ptr(y), total))
        return self._split(x), self._split(y)

    def evaluate(self):
        total = sum(self.counts)
        u, v = np.zeros(total), np.zeros(total)
        _check(self.lib.pv_batch_evaluate(self._h, _ptr(u), _ptr(v), total))
        return self._split(u), self._split(v)

    def failures(self) -> List[str]:
        """Per instance: '' when healthy, else the reason it stopped (it then holds the
        state it had before its failing step)."""
        return [f for f, _ in self.instance_status()]

    def instance_status(self):
        """Per instance: (failure reason or '', RK4 steps completed since load)."""
        k = len(self.counts)
        arr = (ctypes.c_uint * max(1, k))()
        done = (ctypes.c_ulonglong * max(1, k))()
        _check(self.lib.pv_batch_failures(self._h, arr, done, k))
        return [(self.lib.pv_batch_failure_message(arr[i]).decode() if arr[i] else "",
                 int(done[i])) for i in range(k)]

    # -- adaptive runs
    def admit(self, config: InstanceConfig, x, y, g, t_end: float, dt0: float = 1e-3,
              atol: float = 1e-10, rtol: float = 1e-8, min_dt: float = 1e-12,
              max_dt: float = 0.1, precision: Optional[str] = None) -> Optional[int]:
        """Places one DOPRI5 run (in `precision`, default the ensemble's current one); returns
        its slot, or None when it does not fit now."""
        if precision is not None:
            self.set_precision(precision)
        x, y, g = _f64(x), _f64(y), _f64(g)
        slot = ctypes.c_int(-1)
        c = config.c()
        _check(self.lib.pv_batch_admit(self._h, ctypes.byref(c), _ptr(x), _ptr(y), _ptr(g),
                                       len(x), float(t_end), float(dt0), float(atol),
                                       float(rtol), float(min_dt), float(max_dt),
                                       ctypes.byref(slot)))
        return None if slot.value < 0 else slot.value

    def attempt(self, count: int = 1):
        """`count` DOPRI5 attempts of every running run, one synchronisation at the end."""
        if count == 1:
            _check(self.lib.pv_batch_attempt(self._h))
        else:
            _check(self.lib.pv_batch_attempts(self._h, int(count)))

    def poll(self) -> List[SlotState]:
        cap = self.max_instances
        arr = (PvSlotState * cap)()
        count = ctypes.c_size_t(0)
        _check(self.lib.pv_batch_poll(self._h, arr, cap, ctypes.byref(count)))
        out = []
        for k in range(min(count.value, cap)):
            s = arr[k]
            msg = self.lib.pv_batch_failure_message(s.failure).decode() if s.failure else ""
            out.append(SlotState(s.slot, STATES[s.state], s.t, s.dt, s.last_error, s.accepted,
                                 s.rejected, msg, int(s.n)))
        return out

    def fetch(self, slot: int, n: int):
        x, y = np.zeros(n), np.zeros(n)
        _check(self.lib.pv_batch_fetch(self._h, int(slot), _ptr(x), _ptr(y), n))
        return x, y

    def release(self, slot: int):
        _check(self.lib.pv_batch_release(self._h, int(slot)))
