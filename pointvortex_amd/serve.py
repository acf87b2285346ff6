"""A request loop with continuous batching over an adaptive ensemble.

Requests (geometry, N, circulations, horizon, tolerance) arrive over time.  Between DOPRI5
attempts the loop admits, in arrival order, every arrived request that fits the HBM budget, so
new runs join runs already in flight; finished or failed runs leave and free their bytes.  The
budget is the ensemble's device byte limit: a run is admitted only if everything the ensemble
then holds on the device -- its per-vortex state (208 bytes per vortex), the work items and
partial sums of every admitted run (16 bytes per vortex per tile of the run: 64, 128 or 256
vortices by its launch rule), and its tables -- stays within it (`Ensemble.set_byte_limit`).  Each run keeps its state, FSAL
derivative and step-size controller on the device; one attempt advances every running run by
one DOPRI5 attempt (seven stage evaluations).  Between polls the loop runs up to `max_batch`
attempts back to back with one synchronisation: one while an arrived request waits for budget,
never past the next arrival, and never more than a quarter of the attempts the youngest running
run has taken, so that a run finishes at most about a quarter of its attempts before the loop
sees it.  Runs are split into pairs of tiles, each unordered pair computed once
(`Ensemble.set_symmetric`), so a lone large run still spreads over the GPU.  Latency is
measured from arrival to completion.

Precision policy: a request runs in mixed precision when its rtol is at least
`mixed_rtol_threshold` (1e-6 by default), otherwise in FP64.  Both kinds of run live in the
one ensemble, each with its own arithmetic, and advance in the same persistent launches.

Tiles: the serving ensemble takes every run at 64-vortex tiles (`SERVING_RULES`, 1 target per
lane), not the throughput-tuned ensemble table.  Its tail is a few stiff runs left alone on
the GPU, and smaller tiles spread one run over more waves: on the benchmark's request mix
the capacity measured 41.2 solves/s and p99 0.75 s, against 40.3 and 0.97 s at the built-in
128-vortex tiles and 23.4 and 1.70 s with the fixed-step ensembles' table (256-vortex tiles
for most infinite-plane runs).
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import workloads as wl
from .ensemble import Ensemble, InstanceConfig, bytes_per_vortex

SIZE_MAX = 2 ** 64 - 1
SERVING_RULES = [(g, p, SIZE_MAX, 1, 4) for g in ("infinite", "periodic", "disk")
                 for p in ("fp64", "mixed")]


@dataclass
class Request:
    rid: int
    arrival: float            # seconds after the loop starts
    geometry: str
    n: int
    seed: int
    t_end: float
    rtol: float
    atol: float
    dt0: float = 1e-3

    def population(self):
        if self.geometry == "infinite":
            return wl.infinite(self.n, self.seed), InstanceConfig("infinite")
        if self.geometry == "periodic":
            return (wl.periodic(self.n, self.seed, box_length=2 * np.pi),
                    InstanceConfig("periodic", box_length=2 * np.pi))
        return wl.disk(self.n, self.seed), InstanceConfig("disk", disk_radius=1.0)


@dataclass
class Result:
    rid: int
    state: str
    arrival: float
    admitted: float
    finished: float
    accepted: int
    rejected: int
    failure: str
    precision: str
    x: Optional[np.ndarray] = None
    y: Optional[np.ndarray] = None

    @property
    def latency(self):
        return self.finished - self.arrival


@dataclass
class ServerStats:
    results: List[Result] = field(default_factory=list)
    wall: float = 0.0
    attempts: int = 0

    def summary(self):
        """Latency percentiles over the finished runs and the unfinished ones, whose latency
        counts as the time they had waited when the loop stopped (a lower bound): with any
        unfinished run the percentiles are lower bounds too (`latency_lower_bound`).  Requests
        whose arrival time the loop never reached (`not_arrived`) have no latency."""
        done = sum(r.state == "finished" for r in self.results)
        unfinished = sum(r.state == "unfinished" for r in self.results)
        not_arrived = sum(r.state == "not_arrived" for r in self.results)
        lat = np.array([r.latency for r in self.results if r.state in ("finished", "unfinished")])
        return {
            "completed": done,
            "failed": sum(r.state == "failed" for r in self.results),
            "unfinished": unfinished,
            "not_arrived": not_arrived,
            "wall_s": self.wall,
            "solves_per_s": done / self.wall if self.wall > 0 else 0.0,
            "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "latency_lower_bound": unfinished > 0,
            "attempts": self.attempts,
        }


class Server:
    def __init__(self, hbm_budget_bytes: int, max_runs: int = 4096, device: int = 0,
                 mixed_rtol_threshold: float = 1e-6, keep_results: bool = False,
                 symmetric: bool = True, max_batch: int = 16, launch_rules=None):
        self.per_vortex = bytes_per_vortex()
        self.budget = int(hbm_budget_bytes)
        # The ensemble's vortex index is a 32-bit int; its arrays grow as runs are admitted,
        # so the capacity is only a bound (the byte limit binds first on any real budget).
        self.capacity = max(1, min(self.budget // self.per_vortex, (1 << 30) - 1))
        self.used = 0                      # footprint of the admitted runs (estimate)
        self.peak_used = 0
        self.peak_device_bytes = 0         # what the ensemble held on the device, at most
        self.mixed_and_fp64_at_once = False   # runs of both precisions were admitted together
        self.max_runs = max_runs
        self.device = device
        self.threshold = mixed_rtol_threshold
        self.keep = keep_results
        self.symmetric = symmetric
        self.max_batch = max(1, int(max_batch))
        self.launch_rules = SERVING_RULES if launch_rules is None else list(launch_rules)
        self._engine: Optional[Ensemble] = None

    def footprint(self, n: int, geometry: str = "disk", precision: str = "fp64") -> int:
        """HBM bytes a run of n vortices holds while admitted: its state, and its work items
        and partial sums (at the tile the ensemble's launch rules give it)."""
        return n * self.per_vortex + self.engine.instance_bytes(geometry, precision, n,
                                                                 self.symmetric)

    def request_footprint(self, req: Request) -> int:
        return self.footprint(req.n, req.geometry, self.precision_for(req))

    def precision_for(self, req: Request) -> str:
        return "mixed" if req.rtol >= self.threshold else "fp64"

    @property
    def engine(self) -> Ensemble:
        if self._engine is None:
            e = Ensemble(self.capacity, self.max_runs, "fp64", self.device)
            e.set_symmetric(1 if self.symmetric else 0)
            e.set_launch_table(self.launch_rules)
            e.set_byte_limit(self.budget)
            self._engine = e
        return self._engine

    def _track(self):
        self.peak_device_bytes = max(self.peak_device_bytes, self.engine.device_bytes())

    def run(self, requests: List[Request], clock=time.perf_counter,
            max_wall: float = float("inf"), on_result=None) -> ServerStats:
        """Serves `requests` (arrival times relative to the start) until all are done or
        `max_wall` seconds have passed; `on_result(result, stats)` is called as each run
        leaves."""
        eng = self.engine
        pending = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        active: Dict[int, tuple] = {}      # slot -> (request, admitted time, precision)
        tried: Dict[int, int] = {}         # attempts each running run has taken so far
        per_attempt = 0.0                  # running average of the wall time of one attempt
        stats = ServerStats()
        t0 = clock()
        while pending or active:
            now = clock() - t0
            if now > max_wall:
                break
            # Admit, in arrival order, everything that has arrived and fits.
            while pending and pending[0].arrival <= now:
                req = pending[0]
                prec = self.precision_for(req)
                pop, cfg = req.population()
                slot = eng.admit(cfg, pop.x, pop.y, pop.circulation, req.t_end, req.dt0,
                                 req.atol, req.rtol, precision=prec)
                if slot is None:
                    if not active:
                        raise RuntimeError(f"request {req.rid} (N={req.n}) exceeds the budget")
                    break
                self._track()
                self.used += self.request_footprint(req)
                self.peak_used = max(self.peak_used, self.used)
                active[slot] = (req, clock() - t0, prec)
                pending.popleft()
                if len({p for _, _, p in active.values()}) == 2:
                    self.mixed_and_fp64_at_once = True
            if not active:
                wait = pending[0].arrival - (clock() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
                continue
            batch = 1
            if self.max_batch > 1 and not (pending and pending[0].arrival <= now):
                batch = min(self.max_batch, min(tried.get(k, 0) for k in active) // 4)
                if pending and per_attempt > 0.0:
                    batch = min(batch, int((pending[0].arrival - now) / per_attempt))
                batch = max(1, batch)
            started = clock()
            eng.attempt(batch)
            stats.attempts += batch
            self._track()
            elapsed = (clock() - started) / batch
            per_attempt = elapsed if per_attempt == 0.0 else 0.8 * per_attempt + 0.2 * elapsed
            now = clock() - t0
            for s in eng.poll():
                if s.slot not in active:
                    continue
                if s.state == "running":
                    tried[s.slot] = s.accepted + s.rejected
                    continue
                tried.pop(s.slot, None)
                req, admitted, prec = active.pop(s.slot)
                x = y = None
                if self.keep and s.state == "finished":
                    x, y = eng.fetch(s.slot, s.n)
                eng.release(s.slot)
                self.used -= self.request_footprint(req)
                stats.results.append(Result(req.rid, s.state, req.arrival, admitted, now,
                                            s.accepted, s.rejected, s.failure, prec, x, y))
                if on_result is not None:
                    on_result(stats.results[-1], stats)
        stats.wall = clock() - t0
        # Runs still going, or not yet admitted, when max_wall stopped the loop: recorded as
        # unfinished, their latency measured up to now (a lower bound of the real one).
        counts = {s.slot: (s.accepted, s.rejected) for s in eng.poll() if s.slot in active}
        for slot, (req, admitted, prec) in active.items():
            acc, rej = counts.get(slot, (0, 0))
            stats.results.append(Result(req.rid, "unfinished", req.arrival, admitted, stats.wall,
                                        acc, rej, "", prec))
        # Requests whose arrival time the loop never reached are not latencies at all.
        for req in pending:
            state = "unfinished" if req.arrival <= stats.wall else "not_arrived"
            stats.results.append(Result(req.rid, state, req.arrival, float("nan"), stats.wall,
                                        0, 0, "", self.precision_for(req)))
        return stats


def make_requests(count: int, seed: int, rate: float, n_min: int = 64, n_max: int = 131072,
                  t_min: float = 0.1, t_max: float = 1.0, rtol_min: float = 1e-10,
                  rtol_max: float = 1e-6, atol_factor: float = 1e-3,
                  dt0: float = 1e-3) -> List[Request]:
    """Open-loop Poisson arrivals at `rate` per second (all at t=0 when rate is 0); geometry
    uniform over the three; N log-uniform in [n_min, n_max] (even); rtol log-uniform; horizon
    uniform; atol = atol_factor * rtol; initial step dt0; each request seeded by its index.
    Arrival times come from their own random stream, so the requests themselves do not depend
    on the rate."""
    rng = np.random.default_rng(seed)
    gaps = np.random.default_rng([seed, 1]).standard_exponential(count)
    t = 0.0
    out = []
    for k in range(count):
        t += gaps[k] / rate if rate > 0 else 0.0
        geo = ["infinite", "periodic", "disk"][int(rng.integers(0, 3))]
        n = int(np.floor(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        n = max(2, n - n % 2)
        rtol = float(np.exp(rng.uniform(np.log(rtol_min), np.log(rtol_max))))
        horizon = float(rng.uniform(t_min, t_max))
        out.append(Request(k, t, geo, n, seed * 1_000_003 + k, horizon, rtol, atol_factor * rtol,
                           dt0))
    return out
