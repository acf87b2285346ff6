"""Initial-condition files in the reference's format, for the seeded workloads.

The reference's driver reads an initial condition as text lines ``x y circulation`` ('#'
starts a comment, commas count as spaces; src/read.cpp:212-247) and takes the geometry, box
length and disk radius from ``key=value`` fields in the comments (src/read.cpp:248-280).  Its
generator writes a fixed header (initial_conditions/initial_condition.cpp:211-237):

    # PointVortex initial condition
    # geometry=<g> pattern=<p> seed=<s>
    # box_length=<L> disk_radius=<R> infinite_half_width=<h>
    # minimum_separation=<d>
    # x y circulation

followed by one vortex per line at 17 significant digits, so that a file reproduces the
doubles exactly.  `write_initial_condition` writes that header (plus one comment line naming
the workload, which the reader ignores) for the populations of `pointvortex_amd.workloads`,
after the reference generator's checks (initial_condition.cpp:169-196): positive count,
nonzero circulations, periodic vortices inside [-L/2, L/2), disk vortices strictly inside
the wall, and no overwrite unless asked.  `read_initial_condition` parses a file the way the
reference's reader does.

Command line: ``python -m pointvortex_amd.ic {c1,c2,c3,c4,c5} --out DIR [--overwrite]``
writes the configuration's inputs (c4: one file per instance and ``manifest.csv``; c5: the
request table ``manifest.csv`` and the first ``--files`` requests as files).
"""
from __future__ import annotations

import argparse
import csv
import math
import os
from pathlib import Path
from typing import Dict, Optional

import numpy as np

from . import workloads as wl


def _distance(x, y, i, j, geometry, box_length):
    """initial_condition.cpp:33-42 for one pair."""
    dx = float(x[i]) - float(x[j])
    dy = float(y[i]) - float(y[j])
    if geometry == "periodic":
        dx = math.remainder(dx, box_length)
        dy = math.remainder(dy, box_length)
    return math.hypot(dx, dy)


def minimum_separation(pop: wl.Population) -> float:
    """The smallest pair distance (minimum image in the periodic box), as the reference's
    minimumVortexSeparation (initial_condition.cpp:158-167) computes it, for the nearest pair
    a k-d tree finds (the brute-force O(N^2) loop is too slow at N = 65536)."""
    n = pop.n
    if n < 2:
        return math.inf
    from scipy.spatial import cKDTree

    pts = np.stack([pop.x, pop.y], axis=1)
    if pop.geometry == "periodic":
        L = pop.box_length
        tree = cKDTree(np.mod(pts + 0.5 * L, L), boxsize=L)
        d, idx = tree.query(np.mod(pts + 0.5 * L, L), k=2)
    else:
        tree = cKDTree(pts)
        d, idx = tree.query(pts, k=2)
    i = int(np.argmin(d[:, 1]))
    return _distance(pop.x, pop.y, i, int(idx[i, 1]), pop.geometry, pop.box_length)


def validate(pop: wl.Population) -> None:
    """initial_condition.cpp:169-196 (the separation floor is the generator's option, not
    applied here)."""
    if pop.n == 0:
        raise ValueError("vortex count must be positive")
    x, y, g = pop.x, pop.y, pop.circulation
    if not (np.all(np.isfinite(x)) and np.all(np.isfinite(y)) and np.all(np.isfinite(g))):
        raise ValueError("vortex data must be finite")
    if np.any(g == 0.0):
        raise ValueError("initial condition contains an invalid vortex")
    if pop.geometry == "periodic":
        h = 0.5 * pop.box_length
        if np.any((x < -h) | (x >= h) | (y < -h) | (y >= h)):
            raise ValueError("periodic vortex lies outside the fundamental box")
    if pop.geometry == "disk" and np.any(x * x + y * y >= pop.radius * pop.radius):
        raise ValueError("disk vortex lies on or outside the boundary")


def _g17(v: float) -> str:
    return f"{v:.17g}"


def write_initial_condition(path, pop: wl.Population, seed: int, workload: str = "",
                            overwrite: bool = False, infinite_half_width: float = 1.0) -> Path:
    """Writes `pop` in the reference's initial-condition format; returns the path."""
    validate(pop)
    path = Path(path)
    if path.exists() and not overwrite:
        raise FileExistsError(f"refusing to overwrite initial-condition file: {path}")
    path.parent.mkdir(parents=True, exist_ok=True)
    lines = [
        "# PointVortex initial condition",
        f"# geometry={pop.geometry} pattern=random seed={seed}",
        f"# box_length={_g17(pop.box_length)} disk_radius={_g17(pop.radius)} "
        f"infinite_half_width={_g17(infinite_half_width)}",
        f"# minimum_separation={_g17(minimum_separation(pop))}",
    ]
    if workload:
        lines.append(f"# workload={workload.replace(' ', '_')} n={pop.n} "
                     "generator=pointvortex_amd.workloads")
    lines.append("# x y circulation")
    body = "\n".join(f"{_g17(a)} {_g17(b)} {_g17(c)}"
                     for a, b, c in zip(pop.x.tolist(), pop.y.tolist(),
                                        pop.circulation.tolist()))
    tmp = path.with_name(path.name + ".tmp")
    tmp.write_text("\n".join(lines) + "\n" + body + "\n")
    os.replace(tmp, path)
    return path


def read_initial_condition(path) -> wl.Population:
    """src/read.cpp:212-280: vortex lines and the geometry metadata of a file."""
    meta: Dict[str, str] = {}
    xs, ys, gs = [], [], []
    with open(path) as f:
        for number, line in enumerate(f, 1):
            cut = line.find("#")
            if cut >= 0:
                for field in line[cut + 1:].split():
                    if "=" in field:
                        key, value = field.split("=", 1)
                        meta[key] = value
                line = line[:cut]
            fields = line.replace(",", " ").split()
            if not fields:
                continue
            if len(fields) != 3:
                raise ValueError(f"invalid initial condition on line {number}")
            try:
                x, y, g = (float(v) for v in fields)
            except ValueError:
                raise ValueError(f"invalid initial condition on line {number}") from None
            if not (math.isfinite(x) and math.isfinite(y) and math.isfinite(g)):
                raise ValueError(f"invalid initial condition on line {number}")
            xs.append(x)
            ys.append(y)
            gs.append(g)
    if not xs:
        raise ValueError("initial-condition file is empty")
    geometry = meta.get("geometry", "infinite")
    return wl.Population(geometry, np.array(xs), np.array(ys), np.array(gs),
                         box_length=float(meta.get("box_length", 2.0 * math.pi)),
                         radius=float(meta.get("disk_radius", 1.0)))


def write_config(name: str, out, overwrite: bool = False, files: int = 16) -> Path:
    """The inputs of one configuration (c1..c5) under directory `out`."""
    out = Path(out)
    if name == "c1":
        return write_initial_condition(out / "c1_infinite_n1024.ic", wl.config1(), 0,
                                       "infinite N=1024 RK4 dt=1e-3 1000 steps", overwrite)
    if name == "c2":
        return write_initial_condition(out / "c2_periodic_n16384.ic", wl.config2(), 1,
                                       "periodic L=2pi N=16384 DOPRI5", overwrite)
    if name == "c3":
        return write_initial_condition(out / "c3_disk_n65536.ic", wl.config3(), 2,
                                       "disk R=1 N=65536 RK4 dt=1e-4 200 steps", overwrite)
    if name == "c4":
        d = out / "c4_ensemble"
        rows = []
        for k, p in enumerate(wl.ensemble(4096)):
            f = write_initial_condition(d / f"instance_{k:04d}.ic", p, 3,
                                        f"ensemble instance {k}", overwrite)
            rows.append((k, p.geometry, p.n, f.name))
        return _manifest(d / "manifest.csv", ("instance", "geometry", "n", "file"), rows)
    if name == "c5":
        from .serve import make_requests

        d = out / "c5_requests"
        reqs = make_requests(20000, 5, rate=1.0)
        rows = []
        for r in reqs:
            fname = ""
            if r.rid < files:
                pop, _ = r.population()
                fname = write_initial_condition(d / f"request_{r.rid:05d}.ic", pop, r.seed,
                                                f"request {r.rid}", overwrite).name
            rows.append((r.rid, f"{r.arrival:.17g}", r.geometry, r.n, f"{r.t_end:.17g}",
                         f"{r.rtol:.17g}", f"{r.atol:.17g}", fname))
        return _manifest(d / "manifest.csv", ("request", "arrival_at_rate_1", "geometry", "n",
                                              "t_end", "rtol", "atol", "file"), rows)
    raise ValueError(f"unknown configuration {name}")


def _manifest(path: Path, header, rows) -> Path:
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)
    return path


def main(argv: Optional[list] = None) -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("config", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--out", required=True)
    ap.add_argument("--overwrite", action="store_true")
    ap.add_argument("--files", type=int, default=16,
                    help="c5: write the first FILES requests as initial-condition files")
    args = ap.parse_args(argv)
    print(write_config(args.config, args.out, args.overwrite, args.files))


if __name__ == "__main__":
    main()
