"""The persistent RK4 path (all stages of all steps in one cooperative launch) against the
per-stage launches, the oracle and itself; its failure handling; and when it is taken.

Tolerances: the two paths sum each target's sources in different orders, so FP64 positions
agree to rounding (the displacement of a few steps to 1e-12 normwise relative), mixed ones to
the FP32 sum error (2e-6).  Against the oracle the FP64 bar after 10 steps is 1e-12 of the
coordinate scale, widened to 4x the divergence of the oracle's own one-ulp twins where the
population amplifies rounding more than that.
"""
import re

import numpy as np
import pytest

from oracle.pyoracle import spec
from pointvortex_amd import BackendRuntimeError, HipKernel, InvalidArgument, SimParams, VortexSystem
from pointvortex_amd import workloads as wl

pytestmark = pytest.mark.gpu

DT = {"infinite": 1e-3, "periodic": 1e-3, "disk": 1e-4}


def kernel(geometry, precision="fp64", mode=-1, tb=0, wait=0.0):
    p = SimParams(boundaryCondition=geometry, precision=precision)
    if geometry == "periodic":
        p.boxLengthX = p.boxLengthY = 2.0
    k = HipKernel(p)
    k.set_persistent(mode, tb, wait)
    return k


def population(geometry, n, seed):
    """A jittered lattice (nearest neighbours about 1/sqrt(N) apart), so that every pair is
    resolved at the step sizes used here: uniform random points put some pairs so close that
    RK4 at dt = 1e-3 turns them by thousands of radians per step, and then any two summation
    orders part after a step."""
    rng = np.random.default_rng(seed)
    half = 0.9 if geometry == "disk" else 1.0
    m = int(np.ceil(np.sqrt(n * (4.0 / np.pi if geometry == "disk" else 1.0)))) + 1
    cell = 2.0 * half / m
    c = -half + cell * (np.arange(m) + 0.5)
    gx, gy = np.meshgrid(c, c)
    x = gx.ravel() + rng.uniform(-0.25, 0.25, gx.size) * cell
    y = gy.ravel() + rng.uniform(-0.25, 0.25, gy.size) * cell
    keep = np.flatnonzero(x * x + y * y < 0.81) if geometry == "disk" else np.arange(x.size)
    pick = rng.permutation(keep)[:n]
    assert pick.size == n
    g = wl.balanced_circulations(rng, n)
    return wl.Population(geometry, x[pick], y[pick], g,
                         box_length=2.0 if geometry == "periodic" else 2.0 * np.pi)


def advance(k, pop, dt, steps):
    s = VortexSystem(pop.x, pop.y, pop.circulation)
    k.uploadDeviceState(s)
    k.deviceRk4Advance(dt, steps)
    out = s.copy()
    k.downloadDeviceState(out)
    return out


def displacement_error(a, b, pop):
    da = np.concatenate([a.x - pop.x, a.y - pop.y])
    db = np.concatenate([b.x - pop.x, b.y - pop.y])
    return np.linalg.norm(da - db) / max(np.linalg.norm(db), 1e-300)


def test_when_the_persistent_path_is_taken():
    k = kernel("infinite")
    assert k.persistent_blocks(1024) == 64           # 16 targets per block
    assert k.persistent_blocks(4096) == 256
    assert k.persistent_blocks(4097) == 0            # automatic mode stops at 4096
    k.set_persistent(1, 32, 0.0)
    assert k.persistent_blocks(8192) == 256
    assert k.persistent_blocks(65537) == 0
    k.set_persistent(0, 0, 0.0)
    assert k.persistent_blocks(1024) == 0
    f = kernel("infinite")
    f.set_launch(1, 8, 128, 0)                   # a forced shape keeps per-stage launches
    assert f.persistent_blocks(1024) == 0
    for bad in [(2, 0, 0.0), (-2, 0, 0.0), (1, 4, 0.0), (1, 64, 0.0), (1, 8, -1.0),
                (1, 8, float("nan"))]:
        with pytest.raises(InvalidArgument):
            k.set_persistent(*bad)


@pytest.mark.parametrize("n", [2, 300, 1024, 2500])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
def test_persistent_matches_per_stage_launches(geometry, precision, n):
    pop = population(geometry, n, seed=11)
    per = kernel(geometry, precision, mode=1)
    assert per.persistent_blocks(n) == (n + 15) // 16
    a = advance(per, pop, DT[geometry], 3)
    b = advance(kernel(geometry, precision, mode=0), pop, DT[geometry], 3)
    tol = 1e-12 if precision == "fp64" else 2e-6
    assert displacement_error(a, b, pop) <= tol


@pytest.mark.parametrize("tb", [8, 16, 32])
def test_targets_per_block(tb):
    pop = population("infinite", 1000, seed=12)
    k = kernel("infinite", mode=1, tb=tb)
    assert k.persistent_blocks(1000) == (1000 + tb - 1) // tb
    a = advance(k, pop, 1e-3, 4)
    b = advance(kernel("infinite", mode=0), pop, 1e-3, 4)
    assert displacement_error(a, b, pop) <= 1e-12


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
def test_persistent_against_the_oracle(oracle, geometry):
    """Two LDS chunks (N = 2500) and, for the disk, a vortex at the centre (every chunk that
    holds it takes the reference's form).  After 10 steps the oracle's own runs from starts
    one ulp apart differ by up to ~2e-11 here (a fast close pair), so the bar is 1e-12 of the
    coordinate scale or 4x that twin divergence, whichever is larger."""
    pop = population(geometry, 2500, seed=13)
    if geometry == "disk":
        pop.x[1500], pop.y[1500] = 1e-40, 0.0
    k = kernel(geometry, mode=1)
    out = advance(k, pop, DT[geometry], 10)
    s = spec(geometry, box_length=2.0)

    def run(x, y):
        it = oracle.integrator(pop.n)
        x, y = x.copy(), y.copy()
        for _ in range(10):
            it.rk4_step(s, x, y, pop.circulation, DT[geometry])
        return x, y

    x, y = run(pop.x, pop.y)
    rng = np.random.default_rng(0)
    twin = 0.0
    for _ in range(2):
        a, b = run(np.nextafter(pop.x, pop.x + rng.choice([-1.0, 1.0], pop.n)),
                   np.nextafter(pop.y, pop.y + rng.choice([-1.0, 1.0], pop.n)))
        twin = max(twin, np.max(np.abs(a - x)), np.max(np.abs(b - y)))
    scale = np.max(np.abs(np.concatenate([x, y])))
    bar = max(1e-12 * scale, 4.0 * twin)
    assert np.max(np.abs(out.x - x)) <= bar
    assert np.max(np.abs(out.y - y)) <= bar


def test_persistent_is_deterministic_and_step_count_independent():
    pop = wl.config1()
    a = advance(kernel("infinite", mode=1), pop, 1e-3, 7)
    b = advance(kernel("infinite", mode=1), pop, 1e-3, 7)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    k = kernel("infinite", mode=1)
    s = VortexSystem(pop.x, pop.y, pop.circulation)
    k.uploadDeviceState(s)
    for _ in range(7):
        k.deviceRk4Step(1e-3)
    c = s.copy()
    k.downloadDeviceState(c)
    assert np.array_equal(a.x, c.x) and np.array_equal(a.y, c.y)


def failing_run(k, pop, dt, steps):
    s = VortexSystem(pop.x, pop.y, pop.circulation)
    k.uploadDeviceState(s)
    with pytest.raises((BackendRuntimeError, InvalidArgument)) as e:
        k.deviceRk4Advance(dt, steps)
    kept = s.copy()
    k.downloadDeviceState(kept)
    return f"{type(e.value).__name__}: {e.value}", kept


def test_failures_match_the_per_stage_launches():
    # A coincident pair fails in the first evaluation.
    pop = wl.Population("infinite", np.array([0.5, 0.5, 0.1]), np.array([0.2, 0.2, 0.3]),
                        np.array([1.0, -1.0, 1.0]))
    m1, s1 = failing_run(kernel("infinite", mode=1), pop, 1e-3, 5)
    m0, s0 = failing_run(kernel("infinite", mode=0), pop, 1e-3, 5)
    assert m1 == m0 == ("BackendRuntimeError: coincident vortices in singular point-vortex "
                        "kernel (RK4 step 1 of 5)")
    assert np.array_equal(s1.x, pop.x) and np.array_equal(s0.x, pop.x)
    # A dipole driven into the wall with a coarse step leaves the disk some steps later; the
    # state kept is the input of the failing step.
    pop = wl.Population("disk", np.array([0.3, 0.3]), np.array([0.01, -0.01]),
                        np.array([-1.0, 1.0]))
    m1, s1 = failing_run(kernel("disk", mode=1), pop, 0.02, 40)
    m0, s0 = failing_run(kernel("disk", mode=0), pop, 0.02, 40)
    assert m1 == m0 == "InvalidArgument: vortex lies on or outside the disk (RK4 step 9 of 40)"
    assert np.max(np.abs(s1.x - s0.x)) <= 1e-12 and np.max(np.abs(s1.y - s0.y)) <= 1e-12
    assert np.all(s1.x ** 2 + s1.y ** 2 < 1.0)


def test_barrier_wait_bound_ends_the_launch():
    """A wait bound far below one stage makes the barriers time out: the launch ends with an
    error instead of waiting, the state kept is the input of the step that timed out, and the
    handle works after."""
    pop = wl.config1()
    k = kernel("infinite", mode=1, wait=1e-9)
    msg, kept = failing_run(k, pop, 1e-3, 20)
    assert msg.startswith("BackendRuntimeError: grid barrier of the persistent RK4 kernel "
                          "timed out")
    step = int(re.search(r"\(RK4 step (\d+) of 20\)", msg).group(1))
    ref = advance(kernel("infinite", mode=0), pop, 1e-3, step - 1)
    assert np.max(np.abs(kept.x - ref.x)) <= 1e-12 and np.max(np.abs(kept.y - ref.y)) <= 1e-12
    k.set_persistent(1, 0, 0.0)
    a = advance(k, pop, 1e-3, 2)
    b = advance(kernel("infinite", mode=0), pop, 1e-3, 2)
    assert displacement_error(a, b, pop) <= 1e-12


CHILD = """
import sys, numpy as np
sys.path.insert(0, {root!r})
from pointvortex_amd import HipKernel, SimParams, VortexSystem
from pointvortex_amd import workloads as wl
pop = wl.config1()
k = HipKernel(SimParams(boundaryCondition="infinite"))
k.set_persistent(1, 0, 0.0)
s = VortexSystem(pop.x, pop.y, pop.circulation)
k.uploadDeviceState(s)
k.deviceRk4Advance(1e-3, 3)
k.downloadDeviceState(s)
np.save({out!r}, np.concatenate([s.x, s.y]))
print("blocks", k.persistent_blocks(pop.n))
"""


def test_another_process_falls_back_when_the_cooperative_launch_is_refused(tmp_path):
    """This process holds the device's cooperative resources (it has just run a persistent
    launch); a second process asking for the persistent path then either gets it or has its
    cooperative launch refused, in which case it runs the per-stage launches.  Either way it
    finishes with the same result."""
    import subprocess
    import sys
    from pathlib import Path

    pop = wl.config1()
    mine = advance(kernel("infinite", mode=1), pop, 1e-3, 3)
    staged = advance(kernel("infinite", mode=0), pop, 1e-3, 3)
    out = tmp_path / "child.npy"
    root = str(Path(__file__).resolve().parent.parent)
    r = subprocess.run([sys.executable, "-c", CHILD.format(root=root, out=str(out))],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout.strip())          # "blocks 0": refused and fell back; "blocks 64": ran
    child = np.load(out)
    n = pop.n
    c = wl.Population("infinite", child[:n], child[n:], pop.circulation)
    for ref in (mine, staged):
        assert displacement_error(VortexSystem(c.x, c.y, c.circulation), ref, pop) <= 1e-12
