"""Benchmark of the point_vortex_hip velocity backend on one MI355X.

Workload (a "step" is one classical RK4 step = four velocity evaluations with the fused stage
arithmetic): disk R=1, N=65536 point vortices, 1% of them at r in [0.995, 0.9999], circulations
N(0,1)/N, seed 2, dt=1e-4, 200 steps -- the N=64K configuration the headline metric is quoted
on.  Inputs are resident in HBM before the timed region; the timed region is K steps enqueued
back to back with one synchronisation at the end.

Printed: one JSON line.  `value` is FP64 pair-interactions/s (the reference's convention,
N(N-1) per evaluation, src/benchmark.cpp:43).  The `mixed` object holds the mixed-precision run
and its normwise velocity error against the FP64 kernel on the same states.  `roofline` is
the FP64 disk evaluation's algorithmic FLOP rate (36 flops per ordered pair, N(N-1) pairs)
over its average duration from HIP events on the handle's stream.  With the tuned launch table
an evaluation at this N is the symmetric path -- the pair kernel over unordered tile pairs
(each pair computed once, both sides) and its reduction kernel, both inside the events -- so
the hardware executes about half the algorithmic pair count.  Before its own GPU work the
benchmark runs three rocprofv3 counter passes over the same evaluation in child processes
(pointvortex_amd/profiling.py): `roofline.traffic` is the HBM bytes of one evaluation
(FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated) and `roofline.fp64_issue` the share of the
pair kernel's SIMD cycles spent issuing FP64 instructions (--no-counters skips them).
`cpu_baseline` is the reference's own CPU kernel (oracle/_ref, built from the reference
sources) or, if absent, the repo's C restatement, timed on this host's cores on a bounded
sample.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("pair-interactions/s at N=64K (fp64 vs mixed @1e-6 rel); batched solves/s + p99")
FLOPS_PER_PAIR = {"infinite": 12, "disk": 36, "periodic": 12 + 22 * 17}
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3



def log(*a):
    print(*a, file=sys.stderr, flush=True)


def normwise(u, v, ur, vr):
    return float(np.sqrt(np.sum((u - ur) ** 2 + (v - vr) ** 2) / np.sum(ur ** 2 + vr ** 2)))


def run_rk4(pv, pop, precision, steps, warmup, dt):
    k = pv.HipKernel(pv.SimParams(boundaryCondition="disk", diskRadius=1.0, precision=precision))
    s = pv.VortexSystem(pop.x, pop.y, pop.circulation)
    k.uploadDeviceState(s)
    if warmup:
        k.deviceRk4Advance(dt, warmup)
    k.uploadDeviceState(s)  # the timed run starts from the configured initial state
    k.profileBegin()
    t0 = time.perf_counter()
    k.deviceRk4Advance(dt, steps)
    t1 = time.perf_counter()
    kernel_ms, launches = k.profileEnd()
    return k, t1 - t0, kernel_ms, launches


def per_step_error(pv, pop, steps, dt):
    """The mixed RK4 run replayed step by step: after every step the mixed and the FP64
    kernels evaluate the mixed run's current state; returns the normwise error per step."""
    km = pv.HipKernel(pv.SimParams(boundaryCondition="disk", diskRadius=1.0, precision="mixed"))
    kf = pv.HipKernel(pv.SimParams(boundaryCondition="disk", diskRadius=1.0))
    s = pv.VortexSystem(pop.x.copy(), pop.y.copy(), pop.circulation)
    km.uploadDeviceState(s)
    errs = []
    for step in range(steps):
        km.deviceRk4Step(dt)
        km.downloadDeviceState(s)
        kf.uploadDeviceState(s)
        got, ref = km.evaluateDeviceState(), kf.evaluateDeviceState()
        errs.append(normwise(got.x, got.y, ref.x, ref.y))
        if (step + 1) % 50 == 0:
            log(f"[bench] per-step error: {step + 1}/{steps} steps, max so far {max(errs):.3e}")
    km.close()
    kf.close()
    return errs


def per_geometry(pv, n, repeats=3):
    """Evaluation-only pair-interactions/s for the three geometries at N (FP64 and mixed)."""
    from pointvortex_amd import workloads as wl

    out = {}
    pops = {"infinite": wl.infinite(n, 0), "periodic": wl.periodic(n, 1),
            "disk": wl.disk(n, 2)}
    for geo, pop in pops.items():
        row = {}
        ref_vel = None
        for prec in ("fp64", "mixed"):
            k = pv.HipKernel(pv.SimParams(boundaryCondition=geo, boxLengthX=2 * np.pi,
                                          boxLengthY=2 * np.pi, precision=prec))
            k.uploadDeviceState(pv.VortexSystem(pop.x, pop.y, pop.circulation))
            k.profileBegin()
            k.evaluateResident(1)
            k.synchronize()
            one_ms, _ = k.profileEnd()
            # At least `repeats` evaluations and about 30 ms of them, so that the timing does not
            # sit on the clock's ramp after the previous kernel.
            reps = max(repeats, min(50, int(30.0 / max(one_ms, 1e-3))))
            k.profileBegin()
            t0 = time.perf_counter()
            k.evaluateResident(reps)
            k.synchronize()
            dt = time.perf_counter() - t0
            ms, launches = k.profileEnd()
            pairs = reps * n * (n - 1)
            vel = k.evaluateDeviceState()
            if prec == "fp64":
                ref_vel = vel
            peak = PEAK_FP64_TFLOPS if prec == "fp64" else PEAK_FP32_TFLOPS
            kt = ms / 1e3
            tpt, splits, tile, sym = k.resolved_launch(n, n)
            row[prec] = {
                "pairs_per_s": pairs / dt,
                "kernel_pairs_per_s": pairs / kt,
                "kernel_ms_per_eval": ms / launches,
                "launch_shape": [tpt, splits, tile, sym],
                "algorithmic_tflops": pairs * FLOPS_PER_PAIR[geo] / kt / 1e12,
                "frac_of_peak": pairs * FLOPS_PER_PAIR[geo] / kt / 1e12 / peak,
            }
            if prec == "mixed":
                row[prec]["normwise_err_vs_fp64"] = normwise(vel.x, vel.y, ref_vel.x, ref_vel.y)
            k.close()
        row["mixed_speedup"] = row["mixed"]["kernel_pairs_per_s"] / row["fp64"]["kernel_pairs_per_s"]
        out[geo] = row
        out["flops_note"] = (
            "algorithmic_tflops counts the reference's flops per ordered pair (infinite 12, disk 36, "
            "periodic 12 + 22 x 17 image rows) over N(N-1) pairs. The symmetric path (launch_shape[3] "
            "= 1) computes each unordered pair once, and the periodic kernels sum the image rows in "
            "their exact Fourier form, so the executed flops are fewer: a fraction above 1 is a "
            "cheaper evaluation, not a faster FPU.")
        log(f"[bench] per-geometry {geo}: fp64 {row['fp64']['kernel_pairs_per_s']:.3e} "
            f"mixed {row['mixed']['kernel_pairs_per_s']:.3e} pairs/s")
    return out


def ensemble_configs(pops):
    from pointvortex_amd.ensemble import InstanceConfig

    out = []
    for p in pops:
        if p.geometry == "periodic":
            out.append(InstanceConfig("periodic", box_length=p.box_length))
        elif p.geometry == "disk":
            out.append(InstanceConfig("disk", disk_radius=p.radius))
        else:
            out.append(InstanceConfig("infinite"))
    return out


def ensemble_bench(pv, steps, count=4096):
    """The 4096-instance RK4 ensemble (geometry 50/25/25 %, N log-uniform in [32, 4096]),
    one persistent launch per stage, `steps` steps replayed from a captured two-step hipGraph,
    FP64 and mixed.  A solve is one instance's completed `steps`-step run; an instance whose
    state fails (leaves the disk, non-finite velocity) stops there and is not a solve.  Pairs
    are metered as N(N-1) per evaluation of the steps each instance completed."""
    from pointvortex_amd import workloads as wl
    from pointvortex_amd.ensemble import Ensemble

    pops = wl.ensemble(count)
    cfgs = ensemble_configs(pops)
    total = sum(p.n for p in pops)
    args = ([p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
    out = {"instances": count, "vortices": total, "steps": steps, "dt": 1e-3,
           "pairs_per_stage": sum(p.n * (p.n - 1) for p in pops)}
    vel = {}
    for prec in ("fp64", "mixed"):
        e = Ensemble(total, count, prec)
        e.load(cfgs, *args)
        vel[prec] = e.evaluate()
        e.load(cfgs, *args)
        t0 = time.perf_counter()
        e.rk4(1e-3, steps, use_graph=True)
        wall = time.perf_counter() - t0
        status = e.instance_status()
        # Only evaluated pairs are metered: an instance that failed stopped after the steps
        # it completed (four evaluations each; the failing step's partial work is not counted).
        pairs = sum(4 * done * p.n * (p.n - 1) for (_, done), p in zip(status, pops))
        solved = sum(1 for f, _ in status if not f)
        reasons = {}
        for (f, _), p in zip(status, pops):
            if f:
                key = f"{p.geometry}: {f}"
                reasons[key] = reasons.get(key, 0) + 1
        out[prec] = {"wall_s": wall, "solves_per_s": solved / wall, "pairs_per_s": pairs / wall,
                     "completed_instances": solved, "failed_instances": count - solved,
                     "failures": reasons,
                     "decomposition": ("symmetric: pairs of tiles (64, 128 or 256 vortices by "
                                       "the ensemble launch rules), each unordered pair once"
                                       if e.symmetric else
                                       "ordered: 64-target items against all sources")}
        log(f"[bench] ensemble {prec}: {wall:.2f} s, {solved / wall:.1f} solves/s, "
            f"{pairs / wall:.3e} pairs/s, failed {count - solved}: {reasons}")
        e.close()
    out["mixed_speedup"] = out["mixed"]["pairs_per_s"] / out["fp64"]["pairs_per_s"]
    err = {}
    for geo in ("infinite", "periodic", "disk"):
        worst = 0.0
        for k, p in enumerate(pops):
            if p.geometry != geo:
                continue
            worst = max(worst, normwise(vel["mixed"][0][k], vel["mixed"][1][k],
                                        vel["fp64"][0][k], vel["fp64"][1][k]))
        err[geo] = worst
    out["mixed_max_normwise_err_vs_fp64_initial"] = err
    return out


def c2_slice(pv, accepted=40, budget_s=20.0):
    """The periodic configuration (L=2 pi, N=16384, +-1, seed 1) as a slice: DOPRI5 at rtol
    1e-9, atol 1e-12 from dt=1e-3 for `accepted` accepted steps (or `budget_s`), with the
    driver's dipole events after every accepted step (src/main.cpp:327-337): removal below
    1e-3 L and independent reinjection (randomSeed 1234567, src/params.h:64) on the device
    state, and the segment invariants when an event happened.  Pairs are metered as N(N-1)
    per evaluation: six per attempt after the first (FSAL), seven when the derivative was
    invalidated by an event."""
    from pointvortex_amd import workloads as wl

    pop = wl.config2()
    n = pop.n
    k = pv.HipKernel(pv.SimParams(boundaryCondition="periodic", boxLengthX=2 * np.pi,
                                  boxLengthY=2 * np.pi))
    k.uploadDeviceState(pv.VortexSystem(pop.x, pop.y, pop.circulation))
    k.setDipoleRandom(1234567)
    dt, t, steps, attempts, events, fresh = 1e-3, 0.0, 0, 0, 0, 1
    event_s = 0.0
    dts = []
    t0 = time.perf_counter()
    while steps < accepted and time.perf_counter() - t0 < budget_s:
        r = k.deviceDopri5Step(dt, 1e-12, 1e-9, 1e-12, 0.1)
        attempts += 1 + r.rejectedSteps
        steps += 1
        t += r.acceptedTimeStep
        dts.append(r.acceptedTimeStep)
        dt = r.suggestedTimeStep
        s0 = time.perf_counter()
        e = k.processDipoles(1e-3 * 2 * np.pi, "independent")
        if e:
            events += e
            fresh += 1
            k.deviceInvariants()
        event_s += time.perf_counter() - s0
    wall = time.perf_counter() - t0
    evals = 6 * attempts + fresh
    out = {"n": n, "accepted_steps": steps, "attempts": attempts, "t_reached": t,
           "wall_s": wall, "attempts_per_s": attempts / wall,
           "pairs_per_s": evals * n * (n - 1) / wall,
           "dipole_events": events,
           "dipole_event_s_per_step": event_s / max(1, steps),
           "accepted_dt_first_last": [dts[0], dts[-1]] if dts else None}
    log(f"[bench] c2 slice: {steps} accepted steps, {attempts} attempts, {events} dipole "
        f"events in {wall:.2f} s ({out['pairs_per_s']:.3e} pairs/s), t={t:.3e}, "
        f"dt {dts[0]:.2e} -> {dts[-1]:.2e}")
    k.close()
    return out


def c1_whole(pv, steps=1000, dt=1e-3):
    """The infinite-plane configuration whole on the device: N=1024, +-1 (seed 0), RK4
    dt=1e-3 for 1000 steps, FP64 and mixed; 4 evaluations per step, N(N-1) pairs each."""
    from pointvortex_amd import workloads as wl

    pop = wl.config1()
    n = pop.n
    out = {"n": n, "steps": steps, "dt": dt}
    def run(prec, persistent):
        k = pv.HipKernel(pv.SimParams(boundaryCondition="infinite", precision=prec))
        k.set_persistent(persistent, 0, 0.0)
        s = pv.VortexSystem(pop.x, pop.y, pop.circulation)
        k.uploadDeviceState(s)
        k.deviceRk4Advance(dt, 10)
        k.uploadDeviceState(s)
        k.synchronize()
        t0 = time.perf_counter()
        k.deviceRk4Advance(dt, steps)
        k.synchronize()
        wall = time.perf_counter() - t0
        blocks = k.persistent_blocks(n)
        shape = list(k.resolved_launch(n, n))
        k.close()
        return wall, blocks, shape

    for prec in ("fp64", "mixed"):
        # The default path (the persistent RK4 kernel at this N), then the per-stage launches.
        wall, blocks, _ = run(prec, -1)
        staged, _, shape = run(prec, 0)
        out[prec] = {"wall_s": wall, "pairs_per_s": 4 * steps * n * (n - 1) / wall,
                     "persistent_blocks": blocks,
                     "per_stage_launches": {"wall_s": staged, "launch_shape": shape,
                                            "pairs_per_s": 4 * steps * n * (n - 1) / staged}}
    log(f"[bench] c1 whole: fp64 {out['fp64']['wall_s']:.4f} s "
        f"(per-stage launches {out['fp64']['per_stage_launches']['wall_s']:.4f} s), "
        f"mixed {out['mixed']['wall_s']:.4f} s "
        f"(per-stage {out['mixed']['per_stage_launches']['wall_s']:.4f} s) for {steps} steps")
    return out


def cpu_samples(threads):
    """The CPU beside the GPU for the other configurations, each a bounded sample timed with
    the reference's own CPU code (oracle/_ref, else the C restatement): the infinite-plane
    run whole (RK4, 1000 steps at N=1024); 1024 of the 16384 targets of one periodic
    evaluation; one evaluation of 64 ensemble instances stratified by geometry (32 infinite,
    16 periodic, 16 disk) and by N."""
    from oracle import pyoracle
    from pointvortex_amd import workloads as wl

    ref = pyoracle.Reference.available()
    impl = pyoracle.Reference() if ref else pyoracle.Oracle()
    kind = "reference" if ref else "port"
    out = {}
    pop = wl.config1()
    k = pyoracle.spec("infinite")
    t0 = time.perf_counter()
    if ref:
        impl.rk4(k, pop.x, pop.y, pop.circulation, 1e-3, 1000)
    else:
        it = impl.integrator(pop.n)
        x, y = pop.x.copy(), pop.y.copy()
        for _ in range(1000):
            it.rk4_step(k, x, y, pop.circulation, 1e-3)
    wall = time.perf_counter() - t0
    n = pop.n
    out["c1_whole"] = {"value": 4000 * n * (n - 1) / wall, "unit": "pair-interactions/s",
                       "wall_s": wall, "sample": "the whole run: RK4, 1000 steps at N=1024"}
    pop = wl.config2()
    k = pyoracle.spec("periodic", box_length=2 * np.pi, image_layers=8)
    t0 = time.perf_counter()
    impl.evaluate_range(k, pop.x, pop.y, pop.circulation, 0, 1024)
    wall = time.perf_counter() - t0
    out["c2_targets"] = {"value": 1024 * (pop.n - 1) / wall, "unit": "pair-interactions/s",
                         "wall_s": wall,
                         "sample": f"targets 0..1023 of one N={pop.n} periodic evaluation "
                                   "(8 image layers)"}
    pops = wl.ensemble(4096)
    pick = []
    for geo, m in (("infinite", 32), ("periodic", 16), ("disk", 16)):
        idx = sorted((i for i, p in enumerate(pops) if p.geometry == geo),
                     key=lambda i: pops[i].n)
        pick += [idx[int(q)] for q in np.linspace(0, len(idx) - 1, m)]
    pairs, t0 = 0, time.perf_counter()
    for i in pick:
        p = pops[i]
        ks = pyoracle.spec(p.geometry, box_length=getattr(p, "box_length", 1.0),
                           radius=getattr(p, "radius", 1.0))
        impl.evaluate(ks, p.x, p.y, p.circulation)
        pairs += p.n * (p.n - 1)
    wall = time.perf_counter() - t0
    out["c4_instances"] = {"value": pairs / wall, "unit": "pair-interactions/s",
                           "wall_s": wall,
                           "sample": "one evaluation of 64 of the 4096 instances (32 infinite, "
                                     "16 periodic, 16 disk, spread over N)"}
    for v in out.values():
        v.update({"cores": threads, "kind": kind})
    log(f"[bench] CPU samples: " + ", ".join(f"{k} {v['value']:.3e}" for k, v in out.items()))
    return out


SERVING = dict(n_min=64, n_max=2048, t_min=1e-4, t_max=1e-3, rtol_min=1e-10, rtol_max=1e-6)
SERVING_WIDE = dict(n_min=64, n_max=131072, t_min=1e-8, t_max=5e-8, rtol_min=1e-10,
                    rtol_max=1e-6, dt0=1e-10)


def serving_bench(pv, count=128, seed=5, kw=SERVING, budget=4 << 30, max_wall=120.0,
                  name="serving", every=False, top=()):
    """Continuous-batching request loop, scaled down so that it runs in about a minute.
    Default: N log-uniform in [64, 2048], horizon U[1e-4, 1e-3], rtol log-uniform in
    [1e-10, 1e-6], atol = 1e-3 rtol, geometry uniform, 4 GB HBM budget.  The wide variant
    (SERVING_WIDE; its last two requests set to an infinite-plane N=131072 and a disk N=65536
    run, the top of the range the draws may miss) keeps the spec's N range [64, 131072] and its
    256 GB budget with horizons U[1e-8, 5e-8] from dt = 1e-10: a +-1 population of 131072
    vortices has pairs about 1.5e-5 apart, turning at about 7e8 rad/s, so rtol 1e-10 holds
    DOPRI5 to steps near 1e-11 (a few hundred attempts for the largest run).  Capacity is measured first with every request arriving at t=0; the measured run
    then uses Poisson arrivals at 80 % of it."""
    from pointvortex_amd.serve import Server, make_requests

    def progress(phase):
        def cb(result, stats):
            done = len(stats.results)
            if every or done % 16 == 0 or done == count:
                log(f"[bench] {name} {phase}: {done}/{count} done (request {result.rid}, "
                    f"{result.state}, {result.accepted + result.rejected} attempts)")
        return cb

    def requests(rate):
        reqs = make_requests(count, seed, rate=rate, **kw)
        for r, (n, geometry) in zip(reqs[len(reqs) - len(top):], top):   # top of the range
            r.n, r.geometry = n, geometry
        return reqs

    reqs = requests(0.0)
    cap = Server(hbm_budget_bytes=budget, max_runs=count).run(
        reqs, max_wall=max_wall, on_result=progress("capacity")).summary()
    rate = 0.8 * cap["solves_per_s"]
    reqs = requests(rate)
    srv = Server(hbm_budget_bytes=budget, max_runs=count)
    run = srv.run(reqs, max_wall=max_wall, on_result=progress("open loop")).summary()
    log(f"[bench] {name}: capacity {cap['solves_per_s']:.1f} solves/s; at {rate:.1f}/s: "
        f"p50 {run['p50_s']:.3f} s p99 {run['p99_s']:.3f} s")
    ns = [r.n for r in reqs]
    return {"requests": count, "capacity_solves_per_s": cap["solves_per_s"],
            "arrival_rate_per_s": rate, **run, "hbm_budget_bytes": budget,
            "peak_device_bytes": srv.peak_device_bytes, "n_max_drawn": max(ns),
            "requests_n_ge_16384": sum(n >= 16384 for n in ns),
            "workload": (f"N log-uniform [{kw['n_min']},{kw['n_max']}], t U[{kw['t_min']:g},"
                         f"{kw['t_max']:g}], dt0 {kw.get('dt0', 1e-3):g}, rtol log-uniform "
                         f"[{kw['rtol_min']:g},{kw['rtol_max']:g}], atol 1e-3 rtol, geometry "
                         f"uniform, seed {seed}")}


def serving_wide_bench(pv):
    return serving_bench(pv, count=64, seed=7, kw=SERVING_WIDE, budget=256 << 30,
                         max_wall=150.0, name="serving wide", every=True,
                         top=((131072, "infinite"), (65536, "disk")))


def cpu_baseline(pop, budget_s=20.0):
    """Reference CPU disk kernel (src/compute.cpp:286-336) on this host, bounded sample: as many
    full N=65536 evaluations as fit in about budget_s (at least one)."""
    from oracle import pyoracle

    kind = "reference" if pyoracle.Reference.available() else "port"
    impl = pyoracle.Reference() if kind == "reference" else pyoracle.Oracle()
    k = pyoracle.spec("disk", radius=1.0)
    n = len(pop.x)
    t0 = time.perf_counter()
    impl.evaluate(k, pop.x, pop.y, pop.circulation)
    first = time.perf_counter() - t0
    reps = max(0, min(4, int(budget_s / max(first, 1e-3)) - 1))
    t0 = time.perf_counter()
    for _ in range(reps):
        impl.evaluate(k, pop.x, pop.y, pop.circulation)
    total = time.perf_counter() - t0
    evals, secs = (reps, total) if reps else (1, first)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": evals * n * (n - 1) / secs,
        "unit": "pair-interactions/s",
        "cores": threads,
        "kind": kind,
        "sample": f"{evals} FP64 disk evaluation(s) of the N={n} workload state "
                  f"({secs:.1f} s, OpenMP {threads} threads, host nproc {os.cpu_count()})",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes (traffic and FP64 issue null)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the per-geometry table, the ensemble and the serving loop")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--only", choices=["fp64", "geometries", "copy", "ensemble", "serving",
                                       "serving-wide", "counters", "c2",
                                       "c1"],
                    default=None,
                    help="profiling runs: only the FP64 RK4 run, only the per-geometry table, "
                         "or only the 256 MiB copy-kernel calibration")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this backend has no multi-GPU path "
              "(each GPU would run its own share; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    dt = 1e-4
    n = args.n
    counters = None
    if args.only is None and n == 65536 and not args.no_counters:
        # Before this process touches the GPU: rocprofv3 counter passes over the same
        # evaluation, in child processes (HBM traffic, FP64 issue share).
        from pointvortex_amd.profiling import evaluation_counters

        log("[bench] counter passes (rocprofv3, child processes) ...")
        counters = evaluation_counters(n, log=log)

    import pointvortex_amd as pv
    from pointvortex_amd import workloads as wl

    if args.only == "counters":
        # One counter pass's child: the calibration copy, then one RK4 step (four evaluations)
        # of the FP64 disk benchmark population.
        pv.backend.calibration_copy(256 << 20, 1)
        k = pv.HipKernel(pv.SimParams(boundaryCondition="disk", diskRadius=1.0))
        k.uploadDeviceState(pv.VortexSystem(*(lambda p: (p.x, p.y, p.circulation))(wl.disk(n, 2))))
        k.deviceRk4Advance(dt, 1)
        k.synchronize()
        k.close()
        return
    if args.only == "copy":
        ms = pv.backend.calibration_copy(256 << 20, 20)
        log(f"[bench] copy kernel 256 MiB: {ms:.4f} ms ({2 * (256 << 20) / ms / 1e6:.1f} GB/s)")
        return
    if args.only == "geometries":
        print(json.dumps(per_geometry(pv, n)), flush=True)
        return
    if args.only == "ensemble":
        print(json.dumps(ensemble_bench(pv, args.ensemble_steps)), flush=True)
        return
    if args.only == "serving":
        print(json.dumps(serving_bench(pv)), flush=True)
        return
    if args.only == "serving-wide":
        print(json.dumps(serving_wide_bench(pv)), flush=True)
        return
    if args.only == "c2":
        print(json.dumps(c2_slice(pv)), flush=True)
        return
    if args.only == "c1":
        print(json.dumps(c1_whole(pv)), flush=True)
        return
    pop = wl.disk(n, 2)
    pairs_per_step = 4 * n * (n - 1)
    log(f"[bench] disk N={n} RK4 dt={dt} steps={args.steps} warmup={args.warmup}")

    k64, wall64, kms64, launches64 = run_rk4(pv, pop, "fp64", args.steps, args.warmup, dt)
    value = args.steps * pairs_per_step / wall64
    log(f"[bench] fp64: {wall64:.3f} s, {value:.4e} pairs/s, kernel {kms64 / launches64:.3f} ms")
    if args.only == "fp64":
        return
    kmx, wallmx, kmsmx, launchesmx = run_rk4(pv, pop, "mixed", args.steps, args.warmup, dt)
    value_mx = args.steps * pairs_per_step / wallmx
    log(f"[bench] mixed: {wallmx:.3f} s, {value_mx:.4e} pairs/s")

    # Velocity error of the mixed kernel against the FP64 kernel on the mixed run's final state,
    # and on the initial state.
    errs = {}
    fp64_eval = pv.HipKernel(pv.SimParams(boundaryCondition="disk", diskRadius=1.0))
    for state in ("final", "initial"):
        s = pv.VortexSystem(pop.x.copy(), pop.y.copy(), pop.circulation)
        if state == "final":
            kmx.downloadDeviceState(s)
        else:
            kmx.uploadDeviceState(s)
        fp64_eval.uploadDeviceState(s)
        ref = fp64_eval.evaluateDeviceState()
        got = kmx.evaluateDeviceState()
        errs[state] = normwise(got.x, got.y, ref.x, ref.y)
    log(f"[bench] mixed normwise error vs fp64: {errs}")
    step_errs = per_step_error(pv, pop, args.steps, dt)

    avg_ms = kms64 / launches64
    achieved = n * (n - 1) * FLOPS_PER_PAIR["disk"] / (avg_ms / 1e3) / 1e12
    tpt, splits, tile, sym = k64.resolved_launch(n, n)
    kernels = (f"sym_kernel<DiskF64,{tpt},{tile}> + sym_reduce_kernel" if sym else
               f"pair_kernel<DiskF64,{tpt},{tile}> ({splits} source splits)")
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall64 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"disk R=1, N={n} (1% at r in [0.995,0.9999]), G~N(0,1)/N, seed 2, "
                        f"RK4 dt=1e-4, {args.steps} steps",
            "geometry": "disk",
            "n": n,
            "integrator": "rk4",
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": counters["traffic_bytes"] if counters else None,
            "traffic_source": ("rocprofv3 FETCH_SIZE and WRITE_SIZE passes run by this benchmark "
                               "over one RK4 step (pair kernel + reduction per evaluation), each "
                               "scaled by its reading of a 256 MiB copy kernel"),
            "fp64_issue": counters["fp64_issue"] if counters else None,
            "fp64_issue_source": ("rocprofv3 pass of SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 and "
                                  "GRBM_GUI_ACTIVE over the pair kernel: (4 (FMA+MUL+ADD) + 16 "
                                  "TRANS) wave cycles / (4 SIMDs x 256 CUs x cycles)"),
            "counters": counters,
            "kernel": kernels,
            "launch_shape": {"targets_per_thread": tpt, "source_splits": splits, "tile": tile,
                             "symmetric": sym},
            "flops_per_pair": FLOPS_PER_PAIR["disk"],
            "pairs_computed": "N(N-1)/2 unordered pairs" if sym else "N(N-1) ordered pairs",
            "avg_launch_ms": avg_ms,
            "launches": launches64,
        },
        "mixed": {
            "value": value_mx,
            "unit": "pair-interactions/s",
            "dtype": "f64-differences/f32-terms/f64-accumulation",
            "speedup_vs_fp64": value_mx / value,
            "normwise_err_vs_fp64_initial": errs["initial"],
            "normwise_err_vs_fp64_final": errs["final"],
            "per_step_err_vs_fp64": {"steps": len(step_errs), "max": max(step_errs),
                                     "mean": float(np.mean(step_errs)),
                                     "last": step_errs[-1]},
            "avg_launch_ms": kmsmx / launchesmx,
        },
    }
    if not args.no_extra:
        result["per_geometry_n65536"] = per_geometry(pv, n)
        result["infinite_n1024_rk4_whole"] = c1_whole(pv)
        result["periodic_n16384_dopri5_slice"] = c2_slice(pv)
        result["ensemble_4096"] = ensemble_bench(pv, args.ensemble_steps)
        result["serving"] = serving_bench(pv)
        result["serving_wide"] = serving_wide_bench(pv)
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        result["cpu_baseline"] = cpu_baseline(pop)
        log(f"[bench] CPU: {result['cpu_baseline']['value']:.3e} pairs/s")
        if not args.no_extra:
            samples = cpu_samples(result["cpu_baseline"]["cores"])
            result["infinite_n1024_rk4_whole"]["cpu"] = samples["c1_whole"]
            result["periodic_n16384_dopri5_slice"]["cpu"] = samples["c2_targets"]
            result["ensemble_4096"]["cpu"] = samples["c4_instances"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
