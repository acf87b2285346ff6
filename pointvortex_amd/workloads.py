"""Seeded synthetic vortex populations for the benchmark and test workloads.

Draw order (numpy PCG64 via ``numpy.random.default_rng(seed)``), fixed so that every
population can be regenerated bit for bit:

* infinite plane  x ~ U[-1, 1), y ~ U[-1, 1) (x drawn for all vortices first, then y);
  circulations +1 for the first half and -1 for the rest, then shuffled with one
  ``rng.permutation``.
* periodic box    x, y ~ U[-L/2, L/2) in the same order; circulations +-1 balanced, shuffled.
* disk            99 % of vortices: r = R sqrt(U), theta ~ U[0, 2 pi); the last 1 % (at least
  one when N >= 100): r ~ U[0.995 R, 0.9999 R); radii for all vortices first, then angles;
  circulations N(0, 1)/N unless ``balanced`` (+-1 as above).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

TWO_PI = 2.0 * np.pi


@dataclass
class Population:
    geometry: str
    x: np.ndarray
    y: np.ndarray
    circulation: np.ndarray
    box_length: float = 2.0 * np.pi
    radius: float = 1.0

    @property
    def n(self):
        return len(self.x)


def balanced_circulations(rng, n):
    g = np.where(np.arange(n) < n // 2, 1.0, -1.0)
    if n % 2:
        g[-1] = 0.0 if n == 1 else g[-1]
    return rng.permutation(g)


def infinite(n, seed, half_width=1.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-half_width, half_width, n)
    y = rng.uniform(-half_width, half_width, n)
    g = balanced_circulations(rng, n)
    return Population("infinite", x, y, g)


def periodic(n, seed, box_length=TWO_PI):
    if n % 2:
        raise ValueError("a balanced periodic population needs an even N")
    rng = np.random.default_rng(seed)
    half = 0.5 * box_length
    x = rng.uniform(-half, half, n)
    y = rng.uniform(-half, half, n)
    g = balanced_circulations(rng, n)
    return Population("periodic", x, y, g, box_length=box_length)


def disk(n, seed, radius=1.0, wall_fraction=0.01, gaussian=True):
    rng = np.random.default_rng(seed)
    n_wall = int(round(wall_fraction * n)) if wall_fraction > 0 else 0
    n_bulk = n - n_wall
    r_bulk = radius * np.sqrt(rng.uniform(0.0, 1.0, n_bulk))
    r_wall = rng.uniform(0.995 * radius, 0.9999 * radius, n_wall)
    r = np.concatenate([r_bulk, r_wall])
    theta = rng.uniform(0.0, TWO_PI, n)
    x, y = r * np.cos(theta), r * np.sin(theta)
    # Rounding may push a wall vortex onto the wall; pull it inside (never seen in practice).
    rr = x * x + y * y
    bad = rr >= radius * radius
    if np.any(bad):
        s = np.sqrt(0.9999 * radius * radius / rr[bad])
        x[bad] *= s
        y[bad] *= s
    if gaussian:
        g = rng.standard_normal(n) / n
    else:
        g = balanced_circulations(rng, n)
    return Population("disk", x, y, g, radius=radius)


# The workloads of the benchmark configurations.
def config1():
    """Infinite plane, N=1024, seed 0 (RK4 dt=1e-3 to t=1)."""
    return infinite(1024, 0)


def config2():
    """Periodic box L=2 pi, N=16384, seed 1 (DOPRI5 rtol 1e-9, atol 1e-12)."""
    return periodic(16384, 1)


def config3():
    """Disk R=1, N=65536, 1% near the wall, N(0,1)/N circulations, seed 2 (RK4 dt=1e-4)."""
    return disk(65536, 2)


def ensemble_sizes(count=4096, seed=3, n_min=32, n_max=4096):
    """Instance geometries and sizes of the 4096-instance ensemble: 50% infinite, 25% periodic,
    25% disk (in that order before a shuffle), N = floor(exp(U[ln n_min, ln n_max])) rounded
    down to even."""
    rng = np.random.default_rng(seed)
    geos = np.array(["infinite"] * (count // 2) + ["periodic"] * (count // 4) +
                    ["disk"] * (count - count // 2 - count // 4))
    geos = rng.permutation(geos)
    n = np.floor(np.exp(rng.uniform(np.log(n_min), np.log(n_max), count))).astype(np.int64)
    n -= n % 2
    n = np.maximum(n, 2)
    return geos, n


def ensemble(count=4096, seed=3, n_min=32, n_max=4096):
    """Populations of the batched ensemble (infinite [-1,1]^2, periodic L=1, disk R=1)."""
    geos, sizes = ensemble_sizes(count, seed, n_min, n_max)
    out = []
    for k, (geo, n) in enumerate(zip(geos, sizes)):
        s = seed * 1_000_003 + k
        if geo == "infinite":
            out.append(infinite(int(n), s))
        elif geo == "periodic":
            out.append(periodic(int(n), s, box_length=1.0))
        else:
            out.append(disk(int(n), s, radius=1.0, wall_fraction=0.0, gaussian=False))
    return out
