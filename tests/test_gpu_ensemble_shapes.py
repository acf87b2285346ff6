"""Launch shapes of the ensemble's symmetric decomposition: per-instance tiles of 64, 128 or 256
vortices (1, 2 or 4 targets per lane) and chunk sizes, chosen per (geometry, precision,
N-bucket) by the ensemble's launch rules, with one launch per (tile, geometry) group.  Every
shape keeps parity with the oracle; the rules apply per instance; a run's bits depend on its
own rule only; rule changes after a load rebuild the work and the graphs; the byte accounting
follows the tile."""
import numpy as np
import pytest

from pointvortex_amd import InvalidArgument
from pointvortex_amd.autotune import SIZE_MAX
from pointvortex_amd.ensemble import Ensemble
from tests.conftest import fp64_gate, normwise
from tests.test_gpu_ensemble import mixed_population, ospec

pytestmark = pytest.mark.gpu

GEOS = ("infinite", "periodic", "disk")


def load(precision, cfgs, pops, rules=None, symmetric=1):
    e = Ensemble(sum(p.n for p in pops), len(pops) + 1, precision)
    e.set_symmetric(symmetric)
    if rules is not None:
        e.set_launch_table(rules)
    e.load(cfgs, [p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
    return e


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("tpl,chunk", [(1, 1), (1, 8), (2, 4), (4, 1), (4, 16)])
def test_every_shape_keeps_parity(oracle, precision, tpl, chunk):
    """N from 2 to 700: tiles partly filled, one-tile and many-tile instances, every geometry."""
    cfgs, pops = mixed_population(24, 7, n_min=2, n_max=700)
    e = load(precision, cfgs, pops, [(g, precision, SIZE_MAX, tpl, chunk) for g in GEOS])
    us, vs = e.evaluate()
    for c, p, u, v in zip(cfgs, pops, us, vs):
        assert e.resolve_launch(c.geometry, precision, p.n) == (tpl, chunk)
        ur, vr = oracle.evaluate(ospec(c), p.x, p.y, p.circulation)
        gate = (fp64_gate(c.geometry, p.x, p.y, p.circulation, ur, vr) if precision == "fp64"
                else 1e-6)
        assert normwise(u, v, ur, vr) <= gate, (c, p.n, tpl, chunk)


RULES = [("infinite", "fp64", 100, 1, 2), ("infinite", "fp64", 400, 4, 8),
         ("periodic", "fp64", SIZE_MAX, 1, 1), ("disk", "fp64", 300, 4, 1)]


def test_rules_apply_per_instance_and_bucket(oracle):
    """The first rule covering (geometry, precision, n) decides; uncovered instances (large
    disk ones, and every instance of the other precision) take 2 targets per lane and chunks
    of 4.  The resulting mix of groups still evaluates to the oracle's velocities."""
    cfgs, pops = mixed_population(30, 11, n_min=2, n_max=700)
    e = load("fp64", cfgs, pops, RULES)
    for c, p in zip(cfgs, pops):
        want = (2, 4)
        for g, _, n_max, t, k in RULES:
            if g == c.geometry and p.n <= n_max:
                want = (t, k)
                break
        assert e.resolve_launch(c.geometry, "fp64", p.n) == want
        assert e.resolve_launch(c.geometry, "mixed", p.n) == (2, 4)
    us, vs = e.evaluate()
    for c, p, u, v in zip(cfgs, pops, us, vs):
        ur, vr = oracle.evaluate(ospec(c), p.x, p.y, p.circulation)
        assert normwise(u, v, ur, vr) <= fp64_gate(c.geometry, p.x, p.y, p.circulation, ur, vr)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_run_bits_depend_on_its_own_rule_only(precision):
    """An instance advanced alone and among instances of other tiles and geometries (other
    launch groups) ends RK4 with the same bits, graph replays and eager launches alike."""
    cfgs, pops = mixed_population(18, 23, n_min=40, n_max=600)
    rules = [(g, precision, n_max, t, k) for g, _, n_max, t, k in RULES]
    mix = load(precision, cfgs, pops, rules)
    mix.rk4(1e-4, 5)
    xm, ym = mix.read()
    for k in (0, 4, 8, 13):
        alone = load(precision, [cfgs[k]], [pops[k]], rules)
        alone.rk4(1e-4, 5, use_graph=k % 2 == 0)
        xa, ya = alone.read()
        assert np.array_equal(xa[0], xm[k]) and np.array_equal(ya[0], ym[k]), (k, pops[k].n)


def test_new_rules_after_a_load_rebuild_work_and_graphs():
    """Rules set on a loaded ensemble take effect at the next step (captured graphs dropped):
    the run then matches an ensemble that had those rules from its load."""
    cfgs, pops = mixed_population(12, 5, n_min=60, n_max=500)
    rules = [(g, "fp64", SIZE_MAX, 4, 2) for g in GEOS]
    late = load("fp64", cfgs, pops)
    late.rk4(1e-4, 2)               # graphs captured with the built-in shape
    late.set_launch_table(rules)
    late.rk4(1e-4, 4)
    early = load("fp64", cfgs, pops)
    early.rk4(1e-4, 2)
    ref = load("fp64", cfgs, pops, rules)
    ref_x, ref_y = early.read()
    ref.load(cfgs, ref_x, ref_y, [p.circulation for p in pops])
    ref.rk4(1e-4, 4)
    xl, yl = late.read()
    xr, yr = ref.read()
    assert all(np.array_equal(a, b) for a, b in zip(xl, xr))
    assert all(np.array_equal(a, b) for a, b in zip(yl, yr))


@pytest.mark.parametrize("rule", [("disk", "fp64", 10, 3, 4), ("disk", "fp64", 10, 2, 65),
                                  ("disk", "fp64", 10, 2, -1)])
def test_bad_rules_are_refused(rule):
    e = Ensemble(16, 2, "fp64")
    with pytest.raises(InvalidArgument):
        e.set_launch_table([rule])


def test_instance_bytes_follow_the_tile():
    """Work items (16 B), chunk starts (4 B, one per item plus one) and nt x 2n partial doubles
    for nt = ceil(n / (64 tpl)) tiles; what the loaded ensemble holds grows by the same
    partials."""
    n = 1000
    for tpl in (1, 2, 4):
        e = Ensemble(4 * n, 4, "fp64")
        e.set_launch_table([("disk", "fp64", SIZE_MAX, tpl, 4)])
        nt = -(-n // (64 * tpl))
        want = nt * (nt + 1) // 2 * (16 + 4) + 4 + nt * 2 * n * 8
        assert e.instance_bytes("disk", "fp64", n) == want
        assert e.instance_bytes("disk", "mixed", n) == e.instance_bytes("infinite", "fp64", n)
        assert e.instance_bytes("disk", "fp64", n, symmetric=False) == -(-n // 64) * 16
