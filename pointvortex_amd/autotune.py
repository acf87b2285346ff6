"""Autotuner of the pair-kernel launch shape per (geometry, precision, N-bucket).

A shape is (targets per thread, source splits, LDS source tile, symmetric); its blocks per CU
follow from the split count.  The symmetric path computes each unordered pair once over pairs
of tiles; its candidates vary targets per thread (up to 8 on the infinite plane) and the tile.  For every geometry, precision and N-bucket the tuner uploads a seeded
population at the bucket's upper N, evaluates it once with the built-in shape (the
reference result), then for every distinct candidate shape:

* parity gate first: the candidate's velocities must match the built-in shape's to the
  product's own bar -- normwise 1e-12 in FP64; in mixed precision, within 1e-6 of the FP64
  kernel and no more than 1.5x the built-in mixed shape's own distance from it;
* then timing: kernel time from HIP events on the handle's stream, averaged over enough
  launches to cover about 30 ms.  The shapes within 5 % of the fastest are then timed three
  more times, in alternation, and the median of their four timings decides.

Among the shapes that passed the gate, the one with the fewest source splits within
`TIE` (2 %) of the fastest wins: each extra split writes and re-reads one partial sum per
target, so a shape that is not measurably faster should not buy its HBM traffic.  The table
is text, one rule per line:
``geometry precision n_max targets_per_thread source_splits tile symmetric`` ('#' comments;
the last column may be left out and means 0), the format
libpvhip.so reads from $PV_HIP_LAUNCH_TABLE or from launch_table.txt beside the library
(pointvortex_amd.build installs pointvortex_amd/tuning/launch_table.txt there).

Ensembles (``--ensemble``): the symmetric decomposition's tile (64, 128 or 256 vortices: 1, 2
or 4 targets per lane) and chunk size (items a wave takes per counter increment) per
(geometry, precision, N-bucket).  Each bucket is a fixed-step ensemble of equal-size
instances at the bucket's upper N (enough of them for about 2^28 pairs per evaluation, 4 to
4096, so that a step is not dominated by the launches' fixed costs).  Every candidate is
gated first: its velocities against the built-in shape's (2 targets per lane, chunks of 4),
normwise per instance, to the same bars as above (in mixed precision, where the built-in
shape's own error passes 1e-6, no worse than 1.5x it); then RK4 steps replayed from the
ensemble's graph are timed from the loaded state (about 0.1 s per candidate), the close ones
again in alternation.  Among the shapes within `TIE` of the fastest the largest tile wins
(fewest partial sums), then the built-in chunk size.  The per-bucket shapes are then checked
on the 4096-instance mix of the three geometries (mix_check): per precision the table keeps
the per-bucket shapes, the per-bucket shapes the shared kernel takes (no launches of their
own), or none, whichever runs the mix fastest.  The table, one rule per line:
``geometry precision n_max targets_per_lane chunk_items``, is the format libpvhip.so reads
from $PV_HIP_ENSEMBLE_TABLE or from ensemble_table.txt beside the library
(pointvortex_amd.build installs pointvortex_amd/tuning/ensemble_table.txt there).

Usage: python -m pointvortex_amd.autotune [--out FILE] [--geometries ...] [--buckets ...]
       python -m pointvortex_amd.autotune --ensemble [--out FILE] [--geometries ...]
"""
from __future__ import annotations

import argparse
import itertools
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import workloads as wl
from .backend import HipKernel, SimParams, VortexSystem

BUCKETS = (256, 1024, 4096, 16384, 65536, 262144)
TPT = (1, 2, 4)
SYM_TPT = (1, 2, 4, 8)     # 8 where the symmetric kernel has it (infinite plane), else 4
TILES = (128, 256)
SPLITS = (0, 1, 2, 4, 8, 16, 32, 64)
GEOMETRIES = ("infinite", "periodic", "disk")
PRECISIONS = ("fp64", "mixed")
SIZE_MAX = 2 ** 64 - 1
TABLE = Path(__file__).resolve().parent / "tuning" / "launch_table.txt"

FP64_GATE = 1e-12
MIXED_GATE = 1e-6
TIE = 0.02
CONFIRM = 0.05
CONFIRM_ROUNDS = 3


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def normwise(u, v, ur, vr):
    den = math.sqrt(float(np.sum(ur ** 2 + vr ** 2)))
    num = math.sqrt(float(np.sum((u - ur) ** 2 + (v - vr) ** 2)))
    return num / den if den > 0 else num


def population(geometry, n, seed=11):
    n += n % 2
    if geometry == "infinite":
        return wl.infinite(n, seed)
    if geometry == "periodic":
        return wl.periodic(n, seed)
    return wl.disk(n, seed)


def _kernel(geometry, precision, pop):
    k = HipKernel(SimParams(boundaryCondition=geometry, boxLengthX=2 * np.pi,
                            boxLengthY=2 * np.pi, diskRadius=1.0, precision=precision))
    k.set_launch_table([])
    k.uploadDeviceState(VortexSystem(pop.x, pop.y, pop.circulation))
    return k


def _time(k, budget_s=0.03):
    k.evaluateResident(1)
    k.synchronize()
    k.profileBegin()
    k.evaluateResident(1)
    k.synchronize()
    ms, launches = k.profileEnd()
    one = ms / launches
    reps = max(1, min(50, int(budget_s * 1e3 / max(one, 1e-3))))
    k.profileBegin()
    k.evaluateResident(reps)
    k.synchronize()
    ms, launches = k.profileEnd()
    return ms / launches


def tune_one(geometry, precision, n, log=_log):
    """Best gated shape for one (geometry, precision, N); returns a result dict."""
    pop = population(geometry, n)
    n = pop.n
    k = _kernel(geometry, precision, pop)
    base = k.evaluateDeviceState()
    default_shape = k.resolved_launch(n, n)
    default_ms = _time(k)
    ref64, base_err = base, 0.0
    if precision == "mixed":
        k64 = _kernel(geometry, "fp64", pop)
        ref64 = k64.evaluateDeviceState()
        k64.close()
        base_err = normwise(base.x, base.y, ref64.x, ref64.y)
    timed = {default_shape: default_ms}
    seen = {default_shape}
    rejected = 0
    candidates = [(tpt, tile, splits, 0) for tpt, tile, splits in
                  itertools.product(TPT, TILES, SPLITS)]
    candidates += [(tpt, tile, 0, 1) for tpt, tile in itertools.product(SYM_TPT, TILES)]
    for tpt, tile, splits, sym in candidates:
        k.set_launch(tpt, splits, tile, sym)
        shape = k.resolved_launch(n, n)
        if shape in seen:
            continue
        seen.add(shape)
        vel = k.evaluateDeviceState()
        if precision == "fp64":
            ok = normwise(vel.x, vel.y, base.x, base.y) <= FP64_GATE
        else:
            err = normwise(vel.x, vel.y, ref64.x, ref64.y)
            ok = err <= MIXED_GATE and err <= max(1.5 * base_err, 1e-7)
        if not ok:
            rejected += 1
            continue
        timed[shape] = _time(k)
    _confirm(k, timed, n)
    k.close()
    fastest = min(timed.values())
    near = [s for s, ms in timed.items() if ms <= fastest * (1.0 + TIE)]
    # fewest partial sums: a split shape writes splits x 2N, a symmetric one (N/tile) x 2N
    best_shape = min(near, key=lambda s: (partial_slots(s, n), timed[s]))
    best_ms = timed[best_shape]
    res = {"geometry": geometry, "precision": precision, "n": n,
           "default": default_shape, "default_ms": default_ms,
           "best": best_shape, "best_ms": best_ms, "candidates": len(seen),
           "rejected": rejected}
    log(f"[autotune] {geometry:8s} {precision:5s} N={n:7d}: default {default_shape} "
        f"{default_ms:.4f} ms -> {best_shape} {best_ms:.4f} ms "
        f"({len(seen)} shapes, {rejected} failed the gate)")
    return res


def _confirm(k, timed, n):
    """One 30 ms timing per shape carries a few per cent of noise (clocks, the order in which
    the shapes ran), enough to turn the tie rule.  The shapes within CONFIRM of the fastest
    are timed CONFIRM_ROUNDS more times, in alternation, and their medians replace the first
    timings."""
    fastest = min(timed.values())
    close = []
    for s, ms in timed.items():
        if ms <= fastest * (1.0 + CONFIRM):
            k.set_launch(*s)
            if k.resolved_launch(n, n) == s:   # a shape that can be set again as it ran
                close.append(s)
    if len(close) < 2:
        return
    samples = {s: [timed[s]] for s in close}
    for _ in range(CONFIRM_ROUNDS):
        for s in close:
            k.set_launch(*s)
            samples[s].append(_time(k))
    for s in close:
        timed[s] = float(np.median(samples[s]))


def partial_slots(shape, n):
    tpt, splits, _, sym = shape
    return -(-n // sym_tile(tpt)) if sym else (splits if splits > 1 else 0)


def sym_tile(tpt):
    """Vortices per tile of the symmetric kernel: 256 threads x tpt, 128 threads at 8."""
    return (128 if tpt >= 8 else 256) * tpt


def tune(geometries=GEOMETRIES, precisions=PRECISIONS, buckets=BUCKETS, log=_log):
    out = []
    for geometry, precision in itertools.product(geometries, precisions):
        for n in buckets:
            out.append(tune_one(geometry, precision, n, log))
    return out


def table_lines(results):
    """Rules in table order; the largest bucket of each (geometry, precision) also covers
    every larger N."""
    lines = ["# geometry precision n_max targets_per_thread source_splits tile symmetric"]
    groups = {}
    for r in results:
        groups.setdefault((r["geometry"], r["precision"]), []).append(r)
    for (geometry, precision), rs in groups.items():
        rs = sorted(rs, key=lambda r: r["n"])
        for i, r in enumerate(rs):
            n_max = SIZE_MAX if i == len(rs) - 1 else r["n"]
            tpt, splits, tile, sym = r["best"]
            lines.append(f"{geometry} {precision} {n_max} {tpt} {splits} {tile} {sym}  "
                         f"# N={r['n']}: {r['default_ms']:.4f} ms built-in {r['default']}, "
                         f"{r['best_ms']:.4f} ms tuned")
    return lines


def read_table(path):
    """Rules of a table file as tuples (geometry, precision, n_max, tpt, splits, tile,
    symmetric)."""
    rules = []
    for line in Path(path).read_text().splitlines():
        line = line.split("#", 1)[0].split()
        if not line:
            continue
        if len(line) == 6:
            line = line + ["0"]
        if len(line) != 7 or line[6] not in ("0", "1"):
            raise ValueError(f"malformed launch rule: {' '.join(line)}")
        g, p, n, t, s, tl, sy = line
        if g not in GEOMETRIES or p not in PRECISIONS:
            raise ValueError(f"unknown geometry or precision in launch rule: {' '.join(line)}")
        rules.append((g, p, int(n), int(t), int(s), int(tl), int(sy)))
    return rules


# ---------------------------------------------------------------- ensembles

ENSEMBLE_BUCKETS = (64, 128, 256, 512, 1024, 2048, 4096, 16384, 65536)
ENSEMBLE_TPL = (1, 2, 4)
ENSEMBLE_CHUNKS = (1, 2, 4, 8, 16)
ENSEMBLE_DEFAULT = (2, 4)
ENSEMBLE_TABLE = Path(__file__).resolve().parent / "tuning" / "ensemble_table.txt"


def ensemble_population(geometry, n, count, seed=17):
    """`count` instances of n vortices of the 4096-instance ensemble's kinds (infinite
    [-1,1]^2, periodic L=1, disk R=1 without the wall shell)."""
    from .ensemble import InstanceConfig

    n += n % 2
    pops = []
    for k in range(count):
        s = seed * 1_000_003 + k
        if geometry == "infinite":
            pops.append(wl.infinite(n, s))
        elif geometry == "periodic":
            pops.append(wl.periodic(n, s, box_length=1.0))
        else:
            pops.append(wl.disk(n, s, radius=1.0, wall_fraction=0.0, gaussian=False))
    cfg = InstanceConfig(geometry, box_length=1.0, disk_radius=1.0)
    return [cfg] * count, pops


def ensemble_count(n):
    return int(min(4096, max(4, (1 << 28) // (n * n))))


def _ens(precision, cfgs, pops):
    from .ensemble import Ensemble

    e = Ensemble(sum(p.n for p in pops), len(pops), precision)
    e.set_symmetric(1)
    e.set_launch_table([])
    e.load(cfgs, [p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
    return e


def _worst(us, vs, urs, vrs):
    return max(normwise(u, v, ur, vr) for u, v, ur, vr in zip(us, vs, urs, vrs))


def _time_rk4(e, dt=1e-9, budget_s=0.1):
    e.rk4(dt, 2)                     # captures the two-step graph
    t0 = time.perf_counter()
    e.rk4(dt, 2)
    one = (time.perf_counter() - t0) / 2
    steps = 2 * max(1, min(100, int(budget_s / max(one, 1e-5) / 2)))
    t0 = time.perf_counter()
    e.rk4(dt, steps)
    return (time.perf_counter() - t0) / steps * 1e3


def tune_ensemble_one(geometry, precision, n, log=_log):
    count = ensemble_count(n)
    cfgs, pops = ensemble_population(geometry, n, count)
    n = pops[0].n
    e = _ens(precision, cfgs, pops)
    base_u, base_v = e.evaluate()
    ref_u, ref_v, base_err = base_u, base_v, 0.0
    if precision == "mixed":
        e64 = _ens("fp64", cfgs, pops)
        ref_u, ref_v = e64.evaluate()
        e64.close()
        base_err = _worst(base_u, base_v, ref_u, ref_v)
    shapes, rejected = [], 0
    for tpl, chunk in itertools.product(ENSEMBLE_TPL, ENSEMBLE_CHUNKS):
        e.set_launch_table([(geometry, precision, SIZE_MAX, tpl, chunk)])
        u, v = e.evaluate()
        if precision == "fp64":
            ok = _worst(u, v, base_u, base_v) <= FP64_GATE
        else:
            # Where the built-in shape's mixed sums are already past 1e-6 of FP64 (small
            # instances with a vortex close to the wall), a candidate must not be less
            # accurate than it.
            err = _worst(u, v, ref_u, ref_v)
            ok = err <= max(1.5 * base_err, 1e-7) and (err <= MIXED_GATE or
                                                      base_err > MIXED_GATE)
        if ok:
            shapes.append((tpl, chunk))
        else:
            rejected += 1
    if not shapes:
        raise RuntimeError(f"no ensemble shape passed the gate for {geometry} {precision} "
                           f"N={n} (built-in error {base_err:.3g})")
    # Timed from the loaded state at dt = 1e-9: the kernels' work does not depend on dt, and
    # no instance moves far enough to fail (a failed instance stops and would take its work
    # out of the timing).
    e.load(cfgs, [p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
    timed = {}
    for shape in shapes:
        e.set_launch_table([(geometry, precision, SIZE_MAX, *shape)])
        timed[shape] = _time_rk4(e)
    fastest = min(timed.values())
    close = [s for s, ms in timed.items() if ms <= fastest * (1.0 + CONFIRM)]
    if len(close) > 1:
        samples = {s: [timed[s]] for s in close}
        for _ in range(CONFIRM_ROUNDS):
            for s in close:
                e.set_launch_table([(geometry, precision, SIZE_MAX, *s)])
                samples[s].append(_time_rk4(e))
        for s in close:
            timed[s] = float(np.median(samples[s]))
    fails = sum(1 for f in e.failures() if f)
    e.close()
    best = _pick(timed)
    # The best shape that the shared kernel takes: the periodic and disk 4-target items run in
    # launches of their own, which the mix may not repay.
    shared = {s: ms for s, ms in timed.items() if s[0] <= 2 or geometry == "infinite"}
    best_shared = _pick(shared) if shared else ENSEMBLE_DEFAULT
    res = {"geometry": geometry, "precision": precision, "n": n, "count": count,
           "default": ENSEMBLE_DEFAULT, "default_ms": timed.get(ENSEMBLE_DEFAULT),
           "best": best, "best_ms": timed[best], "best_shared": best_shared,
           "candidates": len(timed) + rejected, "rejected": rejected,
           "failed_instances": fails, "builtin_error": base_err}
    dms = res["default_ms"]
    log(f"[autotune] ensemble {geometry:8s} {precision:5s} N={n:6d} x{count:5d}: built-in "
        f"{ENSEMBLE_DEFAULT} {dms if dms is None else round(dms, 4)} ms/step -> {best} "
        f"{timed[best]:.4f} ms ({rejected} failed the gate, {fails} instances failed"
        + (f", built-in mixed error {base_err:.2e})" if precision == "mixed" else ")"))
    return res


def _pick(timed):
    """Among the shapes within TIE of the fastest: the largest tile (fewest partial sums), then
    the built-in chunk size."""
    fastest = min(timed.values())
    near = [s for s, ms in timed.items() if ms <= fastest * (1.0 + TIE)]
    return min(near, key=lambda s: (-s[0], s[1] != ENSEMBLE_DEFAULT[1], timed[s]))


def tune_ensemble(geometries=GEOMETRIES, precisions=PRECISIONS, buckets=ENSEMBLE_BUCKETS,
                  log=_log):
    return [tune_ensemble_one(g, p, n, log)
            for g, p in itertools.product(geometries, precisions) for n in buckets]


def _rules(results, key):
    out = []
    groups = {}
    for r in results:
        groups.setdefault((r["geometry"], r["precision"]), []).append(r)
    for (geometry, precision), rs in groups.items():
        rs = sorted(rs, key=lambda r: r["n"])
        for i, r in enumerate(rs):
            n_max = SIZE_MAX if i == len(rs) - 1 else r["n"]
            out.append((geometry, precision, n_max, *r[key]))
    return out


def mix_check(results, steps=40, log=_log):
    """The 4096-instance ensemble (pointvortex_amd.workloads.ensemble: the three geometries,
    N log-uniform in [32, 4096]) per precision under three tables: the per-bucket shapes
    ("best"), the per-bucket shapes the shared kernel takes ("best_shared": one launch), and
    none (the built-in shape).  RK4 steps from graph replays, the three alternated twice; the
    fastest decides that precision's rules.  Returns {precision: (choice, {choice: ms})}."""
    from .ensemble import Ensemble, InstanceConfig

    pops = wl.ensemble(4096)
    cfgs = [InstanceConfig(p.geometry, box_length=getattr(p, "box_length", 2.0),
                           disk_radius=getattr(p, "radius", 1.0)) for p in pops]
    args = ([p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
    total = sum(p.n for p in pops)
    out = {}
    for precision in sorted({r["precision"] for r in results}):
        mine = [r for r in results if r["precision"] == precision]
        tables = {"best": _rules(mine, "best"), "best_shared": _rules(mine, "best_shared"),
                  "built-in": []}
        e = Ensemble(total, len(pops), precision)
        times = {k: [] for k in tables}
        for _ in range(2):
            for name, rules in tables.items():
                e.set_launch_table(rules)
                e.load(cfgs, *args)
                e.rk4(1e-3, 2)
                e.load(cfgs, *args)
                t0 = time.perf_counter()
                e.rk4(1e-3, steps)
                times[name].append((time.perf_counter() - t0) / steps * 1e3)
        e.close()
        ms = {k: min(v) for k, v in times.items()}
        choice = min(ms, key=ms.get)
        out[precision] = (choice, ms)
        log(f"[autotune] ensemble mix {precision}: " +
            ", ".join(f"{k} {v:.3f} ms/step" for k, v in ms.items()) + f" -> {choice}")
    return out


def ensemble_table_lines(results, mix=None):
    """Rules in table order.  With `mix` (from mix_check), each precision takes the shapes its
    mix choice names (none for the built-in shape)."""
    lines = ["# geometry precision n_max targets_per_lane chunk_items"]
    if mix:
        for precision, (choice, ms) in sorted(mix.items()):
            lines.append(f"# {precision} mix (4096 instances): " +
                         ", ".join(f"{k} {v:.3f} ms/step" for k, v in ms.items()) +
                         f"; rules below: {choice}")
    groups = {}
    for r in results:
        groups.setdefault((r["geometry"], r["precision"]), []).append(r)
    for (geometry, precision), rs in groups.items():
        key = "best"
        if mix and precision in mix:
            key = mix[precision][0]
            if key == "built-in":
                continue
        rs = sorted(rs, key=lambda r: r["n"])
        for i, r in enumerate(rs):
            n_max = SIZE_MAX if i == len(rs) - 1 else r["n"]
            tpl, chunk = r[key]
            dms = r["default_ms"]
            note = f"{dms:.4f} ms/step built-in {r['default']}" if dms else "built-in gated out"
            lines.append(f"{geometry} {precision} {n_max} {tpl} {chunk}  # N={r['n']} "
                         f"x{r['count']}: {note}, {r['best_ms']:.4f} tuned")
    return lines


def read_ensemble_table(path):
    """Rules of an ensemble table file as tuples (geometry, precision, n_max,
    targets_per_lane, chunk_items)."""
    rules = []
    for line in Path(path).read_text().splitlines():
        line = line.split("#", 1)[0].split()
        if not line:
            continue
        if len(line) != 5:
            raise ValueError(f"malformed ensemble launch rule: {' '.join(line)}")
        g, p, n, t, c = line
        if g not in GEOMETRIES or p not in PRECISIONS:
            raise ValueError(f"unknown geometry or precision in ensemble launch rule: "
                             f"{' '.join(line)}")
        rules.append((g, p, int(n), int(t), int(c)))
    return rules


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--ensemble", action="store_true",
                    help="tune the ensemble's symmetric decomposition instead")
    ap.add_argument("--out", default=None)
    ap.add_argument("--geometries", nargs="+", default=list(GEOMETRIES))
    ap.add_argument("--precisions", nargs="+", default=list(PRECISIONS))
    ap.add_argument("--buckets", nargs="+", type=int, default=None)
    args = ap.parse_args()
    t0 = time.perf_counter()
    if args.ensemble:
        args.out = args.out or str(ENSEMBLE_TABLE)
        results = tune_ensemble(args.geometries, args.precisions,
                                args.buckets or list(ENSEMBLE_BUCKETS))
        mix = mix_check(results) if set(args.geometries) == set(GEOMETRIES) else None
        lines = ensemble_table_lines(results, mix)
    else:
        args.out = args.out or str(TABLE)
        results = tune(args.geometries, args.precisions, args.buckets or list(BUCKETS))
        lines = table_lines(results)
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    Path(args.out).write_text("\n".join(lines) + "\n")
    _log(f"[autotune] wrote {args.out} ({len(lines) - 1} rules, {time.perf_counter() - t0:.0f} s)")


if __name__ == "__main__":
    main()
