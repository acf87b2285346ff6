"""Launch shapes of the pair kernels: every (targets per thread, splits, tile) variant keeps
parity with the oracle; launch-table precedence and loading; the autotuner's gate."""
import os

import numpy as np
import pytest

from oracle.pyoracle import spec
from pointvortex_amd import HipKernel, InvalidArgument, SimParams, VortexSystem
from pointvortex_amd import autotune
from pointvortex_amd import workloads as wl
from tests.conftest import fp64_gate, normwise

pytestmark = pytest.mark.gpu


def setup(geometry, precision, n, seed=31):
    if geometry == "infinite":
        pop = wl.infinite(n, seed)
    elif geometry == "periodic":
        pop = wl.periodic(n, seed, box_length=2.0)
    else:
        pop = wl.disk(n, seed)
    k = HipKernel(SimParams(boundaryCondition=geometry, precision=precision, boxLengthX=2.0,
                            boxLengthY=2.0))
    return pop, k


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("tpt,splits,tile", [(1, 0, 128), (2, 5, 128), (4, 0, 256),
                                             (1, 2, 256), (4, 9, 128)])
def test_shape_variants_keep_parity(oracle, geometry, precision, tpt, splits, tile):
    pop, k = setup(geometry, precision, 1500)
    k.set_launch(tpt, splits, tile)
    assert k.resolved_launch(pop.n, pop.n)[0] == tpt
    assert k.resolved_launch(pop.n, pop.n)[2] == tile
    vel = k.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
    ur, vr = oracle.evaluate(spec(geometry, box_length=2.0), pop.x, pop.y, pop.circulation)
    gate = (fp64_gate(geometry, pop.x, pop.y, pop.circulation, ur, vr) if precision == "fp64"
            else 1e-6)
    assert normwise(vel.x, vel.y, ur, vr) <= gate


def test_table_precedence():
    pop, k = setup("disk", "fp64", 5000)
    k.set_launch_table([])
    default = k.resolved_launch(pop.n, pop.n)
    k.set_launch_table([("infinite", "fp64", 10**9, 4, 3, 128),     # other geometry
                        ("disk", "mixed", 10**9, 4, 3, 128),        # other precision
                        ("disk", "fp64", 1000, 4, 3, 128),          # N too large for it
                        ("disk", "fp64", 10**6, 1, 2, 128),         # this one
                        ("disk", "fp64", 10**9, 4, 7, 256)])
    assert k.resolved_launch(pop.n, pop.n) == (1, 2, 128, 0)
    k.set_launch(4, 0, 256)                                          # explicit override wins
    assert k.resolved_launch(pop.n, pop.n)[0] == 4 and k.resolved_launch(pop.n, pop.n)[2] == 256
    k.set_launch(0, 0, 0)
    assert k.resolved_launch(pop.n, pop.n) == (1, 2, 128, 0)
    k.set_launch_table([("disk", "fp64", 10**6, 2, 0, 256, 1)])
    assert k.resolved_launch(pop.n, pop.n) == (2, 0, 256, 1)        # symmetric rule
    assert k.resolved_launch(100, pop.n)[3] == 0                     # subsets: split path
    k.set_launch(0, 0, 0, 0)
    assert k.resolved_launch(pop.n, pop.n)[3] == 0                   # forced off
    k.set_launch_table([])
    assert k.resolved_launch(pop.n, pop.n) == default
    with pytest.raises(InvalidArgument, match="tile"):
        k.set_launch_table([("disk", "fp64", 10, 1, 0, 64)])


def test_table_from_environment(tmp_path):
    table = tmp_path / "launch_table.txt"
    table.write_text("# comment\n\ndisk fp64 100000 4 3 128  # a rule\n")
    bad = tmp_path / "bad.txt"
    bad.write_text("disk fp64 100000 4\n")
    old = os.environ.get("PV_HIP_LAUNCH_TABLE")
    try:
        os.environ["PV_HIP_LAUNCH_TABLE"] = str(table)
        k = HipKernel(SimParams(boundaryCondition="disk"))
        assert k.resolved_launch(5000, 5000) == (4, 3, 128, 0)
        os.environ["PV_HIP_LAUNCH_TABLE"] = str(bad)
        with pytest.raises(InvalidArgument, match="malformed"):
            HipKernel(SimParams(boundaryCondition="disk"))
        os.environ["PV_HIP_LAUNCH_TABLE"] = str(tmp_path / "missing.txt")
        with pytest.raises(InvalidArgument, match="cannot read"):
            HipKernel(SimParams(boundaryCondition="disk"))
    finally:
        if old is None:
            os.environ.pop("PV_HIP_LAUNCH_TABLE", None)
        else:
            os.environ["PV_HIP_LAUNCH_TABLE"] = old


@pytest.mark.parametrize("geometry,precision", [("infinite", "fp64"), ("disk", "mixed")])
def test_autotuner_choice_keeps_parity(oracle, geometry, precision):
    res = autotune.tune_one(geometry, precision, 2048, log=lambda *a: None)
    assert res["best_ms"] <= res["default_ms"] * (1.0 + autotune.TIE)
    tpt, splits, tile, sym = res["best"]
    pop = autotune.population(geometry, 2048)
    k = HipKernel(SimParams(boundaryCondition=geometry, precision=precision,
                            boxLengthX=2 * np.pi, boxLengthY=2 * np.pi))
    k.set_launch(tpt, splits, tile, sym)
    vel = k.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
    ur, vr = oracle.evaluate(spec(geometry, box_length=2 * np.pi), pop.x, pop.y, pop.circulation)
    gate = (fp64_gate(geometry, pop.x, pop.y, pop.circulation, ur, vr) if precision == "fp64"
            else 1e-6)
    assert normwise(vel.x, vel.y, ur, vr) <= gate


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
@pytest.mark.parametrize("n", [2, 700, 1500, 5000])
@pytest.mark.parametrize("tpt,tile", [(1, 128), (2, 256), (4, 128)])
def test_symmetric_path_keeps_parity(oracle, geometry, n, tpt, tile):
    """Each unordered pair once over pairs of tiles (N not a multiple of the tile included)."""
    pop, k = setup(geometry, "fp64", n, seed=5 + n)
    k.set_launch(tpt, 0, tile, 1)
    assert k.resolved_launch(pop.n, pop.n) == (tpt, 0, tile, 1)
    vel = k.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
    ur, vr = oracle.evaluate(spec(geometry, box_length=2.0), pop.x, pop.y, pop.circulation)
    assert normwise(vel.x, vel.y, ur, vr) <= fp64_gate(geometry, pop.x, pop.y, pop.circulation,
                                                       ur, vr)


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
@pytest.mark.parametrize("n", [2, 700, 5000])
@pytest.mark.parametrize("tpt,tile", [(1, 128), (2, 256), (4, 128)])
def test_symmetric_mixed_keeps_parity(oracle, geometry, n, tpt, tile):
    """Mixed precision over unordered pairs: within 1e-6 of the oracle, and no less accurate
    than the split path's mixed sums (the reaction is an FP32 sum of at most 64 x 4 terms, the
    same class as the split path's per-tile FP32 sums)."""
    pop, k = setup(geometry, "mixed", n, seed=9 + n)
    s = VortexSystem(pop.x, pop.y, pop.circulation)
    k.set_launch(tpt, 0, tile, 1)
    assert k.resolved_launch(pop.n, pop.n) == (tpt, 0, tile, 1)
    vel = k.evaluate(s)
    k.set_launch(tpt, 4, tile, 0)
    split = k.evaluate(s)
    ur, vr = oracle.evaluate(spec(geometry, box_length=2.0), pop.x, pop.y, pop.circulation)
    err, err_split = normwise(vel.x, vel.y, ur, vr), normwise(split.x, split.y, ur, vr)
    assert err <= 1e-6
    assert err <= max(3.0 * err_split, 1e-7), (err, err_split)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n", [2, 700, 1500, 5000])
@pytest.mark.parametrize("tile", [128, 256])
def test_symmetric_eight_targets_per_thread(oracle, precision, n, tile):
    """8 targets per thread on the infinite plane: 128-thread blocks over 1024-vortex tiles, the
    product tree over 8 denominators.  FP64 within the rounding gate of the oracle; mixed within
    1e-6 and no less accurate than the 4-target symmetric kernel's sums."""
    pop, k = setup("infinite", precision, n, seed=21 + n)
    s = VortexSystem(pop.x, pop.y, pop.circulation)
    k.set_launch(8, 0, tile, 1)
    assert k.resolved_launch(pop.n, pop.n) == (8, 0, tile, 1)
    vel = k.evaluate(s)
    ur, vr = oracle.evaluate(spec("infinite"), pop.x, pop.y, pop.circulation)
    err = normwise(vel.x, vel.y, ur, vr)
    if precision == "fp64":
        assert err <= fp64_gate("infinite", pop.x, pop.y, pop.circulation, ur, vr)
    else:
        k.set_launch(4, 0, tile, 1)
        four = k.evaluate(s)
        assert err <= 1e-6
        assert err <= max(3.0 * normwise(four.x, four.y, ur, vr), 1e-7)


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
def test_eight_targets_per_thread_elsewhere_run_at_four(geometry):
    """Only the infinite plane's symmetric kernel has 8 targets per thread: a request for 8 on
    the split path, or on the other geometries, resolves to (and runs at) 4."""
    pop, k = setup(geometry, "fp64", 700, seed=3)
    k.set_launch(8, 3, 128, 0)
    assert k.resolved_launch(pop.n, pop.n) == (4, 3, 128, 0)
    if geometry != "infinite":
        k.set_launch(8, 0, 128, 1)
        assert k.resolved_launch(pop.n, pop.n) == (4, 0, 128, 1)
        vel = k.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
        k.set_launch(4, 0, 128, 1)
        four = k.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
        assert np.array_equal(vel.x, four.x) and np.array_equal(vel.y, four.y)


@pytest.mark.parametrize("pos", [(2e-160, 1e-161), (-3e-31, 2e-31), (6e-13, -4e-13),
                                 (8e-13, 5e-13), (3e-7, 1e-7)])
def test_symmetric_mixed_near_centre_disk_source(oracle, pos):
    """A near-centre vortex (its image a far, finite offset) in a middle tile, so that it is a
    source for earlier tiles and a target for later ones.  Below 2^-40 R (about 9.1e-13) its
    blocks take the direct form of i's image term; above it, the shared reciprocal with
    w = 1/|x|^2 up to about 1e24.  One such vortex only: two of them would sit closer than
    FP32 can square."""
    pop, k = setup("disk", "mixed", 1800, seed=73)
    x, y = pop.x.copy(), pop.y.copy()
    x[900], y[900] = pos
    k.set_launch(2, 0, 128, 1)
    vel = k.evaluate(VortexSystem(x, y, pop.circulation))
    assert np.all(np.isfinite(vel.x)) and np.all(np.isfinite(vel.y))
    ur, vr = oracle.evaluate(spec("disk"), x, y, pop.circulation)
    assert normwise(vel.x, vel.y, ur, vr) <= 1e-6
    near = np.hypot(vel.x[900] - ur[900], vel.y[900] - vr[900]) / np.hypot(ur[900], vr[900])
    assert near <= 1e-5


@pytest.mark.parametrize("geometry", ["infinite", "disk"])
@pytest.mark.parametrize("tpt,tile", [(1, 128), (2, 256), (4, 128)])
def test_symmetric_fp64_close_pair_across_tiles(oracle, geometry, tpt, tile):
    """A close pair with its two vortices in different tiles of the symmetric FP64 path, so
    that it is evaluated in an off-diagonal block (the paired ring steps at 1-4 denominators
    per step, the single steps at 4 targets on the disk).  On the infinite plane the pair sits
    at the origin, down to 1e-150 apart: its denominator takes the ring's product tree out of
    range and the wave falls back to one reciprocal per denominator.  In the disk (the pair at
    (0.3, 0.1), where FP64 positions resolve about 1e-16) down to 1e-15.  The pair within 1e-10
    of the oracle per vortex, the rest within normwise 1e-12."""
    pop, k = setup(geometry, "fp64", 1500, seed=41)
    k.set_launch(tpt, 0, tile, 1)
    assert k.resolved_launch(pop.n, pop.n) == (tpt, 0, tile, 1)
    i, j = 100, 1400                      # different tiles at 256, 512 and 1024 targets
    if geometry == "infinite":
        x0, y0, seps = 0.0, 0.0, (1e-2, 1e-6, 1e-12, 1e-50, 1e-150)
    else:
        x0, y0, seps = 0.3, 0.1, (1e-2, 1e-6, 1e-12, 1e-15)
    rest = np.ones(pop.n, dtype=bool)
    rest[[i, j]] = False
    for sep in seps:
        x, y = pop.x.copy(), pop.y.copy()
        x[i], y[i] = x0, y0
        x[j], y[j] = x0 + 0.8 * sep, y0 + 0.6 * sep
        vel = k.evaluate(VortexSystem(x, y, pop.circulation))
        ur, vr = oracle.evaluate(spec(geometry), x, y, pop.circulation)
        for v in (i, j):
            err = np.hypot(vel.x[v] - ur[v], vel.y[v] - vr[v])
            assert err <= 1e-10 * np.hypot(ur[v], vr[v]), (sep, v)
        assert normwise(vel.x[rest], vel.y[rest], ur[rest], vr[rest]) <= 1e-12, sep


def test_symmetric_path_near_centre_disk_source(oracle):
    pop, k = setup("disk", "fp64", 1800, seed=71)
    x, y = pop.x.copy(), pop.y.copy()
    x[1300], y[1300] = 2e-160, 1e-161
    k.set_launch(2, 0, 128, 1)
    vel = k.evaluate(VortexSystem(x, y, pop.circulation))
    ur, vr = oracle.evaluate(spec("disk"), x, y, pop.circulation)
    assert normwise(vel.x, vel.y, ur, vr) <= fp64_gate("disk", x, y, pop.circulation, ur, vr)


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
def test_symmetric_path_in_rk4_and_fallbacks(oracle, geometry):
    """RK4 steps through the symmetric path agree with the split path; target subsets fall
    back to the split path; repeats are bit-identical."""
    pop, ks = setup(geometry, "fp64", 3000, seed=12)
    _, kp = setup(geometry, "fp64", 3000, seed=12)
    ks.set_launch(2, 0, 256, 1)
    kp.set_launch(2, 0, 256, 0)
    s = VortexSystem(pop.x, pop.y, pop.circulation)
    a, b = ks.evaluate(s), ks.evaluate(s)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    for kern in (ks, kp):
        kern.uploadDeviceState(s)
        for _ in range(3):
            kern.deviceRk4Step(1e-4)
    es, ep = s.copy(), s.copy()
    ks.downloadDeviceState(es)
    kp.downloadDeviceState(ep)
    scale = np.max(np.abs(np.concatenate([ep.x, ep.y])))
    assert np.max(np.abs(es.x - ep.x)) <= 1e-12 * scale
    assert np.max(np.abs(es.y - ep.y)) <= 1e-12 * scale
    assert ks.resolved_launch(100, 3000)[3] == 0
    sub = ks.evaluateRange(pop.x, pop.y, pop.circulation, 100, 400)
    ur, vr = oracle.evaluate_range(spec(geometry, box_length=2.0), pop.x, pop.y,
                                   pop.circulation, 100, 400)
    assert normwise(sub.x[100:400], sub.y[100:400], ur[100:400], vr[100:400]) <= 1e-12
    _, km = setup(geometry, "mixed", 3000, seed=12)
    km.set_launch(2, 0, 256, 1)
    assert km.resolved_launch(3000, 3000)[3] == 1
    assert km.resolved_launch(100, 3000)[3] == 0


def test_symmetric_path_respects_the_partial_limit():
    """A forced symmetric shape runs while its partial sums (nt x 2N doubles) stay under 8 GiB
    and falls back to the split path above (resolution only; nothing is allocated)."""
    _, k = setup("disk", "fp64", 64)
    k.set_launch(4, 0, 128, 1)
    assert k.resolved_launch(262144, 262144)[3] == 1          # 256 rows x 2N: 1.1 GB
    assert k.resolved_launch(10 ** 6, 10 ** 6)[3] == 0        # 977 rows x 2N: 15.6 GB
    assert k.resolved_launch(1000, 10 ** 6)[3] == 0           # target subsets never
