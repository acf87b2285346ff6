"""The benchmark configurations at their real sizes, against the oracle and the reference build.

* Periodic box, L = 2 pi, N = 16384, +-1 circulations, seed 1, 8 image layers: velocities on
  target subsets (FP64 and mixed), the dipole event at removal distance 1e-3 L with independent
  reinjection (bit for bit against the reference's DipoleManager::process), and one accepted
  DOPRI5 step (rtol 1e-9, atol 1e-12) against the oracle's integrator.
* The 4096-instance ensemble (50 % infinite [-1, 1]^2, 25 % periodic L = 1, 25 % disk R = 1,
  N log-uniform in [32, 4096], seed 3): the whole ensemble evaluated on the device, a sample of
  64 instances stratified by geometry and N against the oracle; and the 500-step RK4 run at
  dt = 1e-3, whose failing disk instances the oracle must fail the same way, at a step inside
  the window its own one-ulp twins give.
"""
import numpy as np
import pytest

from oracle.pyoracle import spec
from pointvortex_amd import HipKernel, SimParams, VortexSystem
from pointvortex_amd import workloads as wl
from pointvortex_amd.ensemble import Ensemble, InstanceConfig
from tests.conftest import disk_truth, fp64_gate, normwise

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * np.pi


# ------------------------------------------------------------------ periodic N=16384


@pytest.fixture(scope="module")
def c2():
    return wl.config2()


def c2_kernel(precision="fp64"):
    return HipKernel(SimParams(boundaryCondition="periodic", boxLengthX=TWO_PI,
                               boxLengthY=TWO_PI, periodicImageLayers=8, precision=precision))


C2_SPEC = spec("periodic", box_length=TWO_PI, image_layers=8)


@pytest.mark.parametrize("precision,bar", [("fp64", 1e-12), ("mixed", 1e-6)])
def test_c2_velocity_on_target_subsets(oracle, c2, precision, bar):
    """The whole N=16384 evaluation on the device (the tuned symmetric path) against the
    oracle's sums for the first, a middle and the last 256 targets over all sources."""
    n = c2.n
    vel = c2_kernel(precision).evaluate(VortexSystem(c2.x, c2.y, c2.circulation))
    for b, e in [(0, 256), (n // 2, n // 2 + 256), (n - 256, n)]:
        ur, vr = oracle.evaluate_range(C2_SPEC, c2.x, c2.y, c2.circulation, b, e)
        assert normwise(vel.x[b:e], vel.y[b:e], ur[b:e], vr[b:e]) <= bar, (b, e)


def test_c2_dipole_event_matches_the_reference(reference, c2):
    """The driver's dipole event on the configuration's initial state (src/main.cpp:327-337):
    removal below 1e-3 L, independent reinjection, generator seeded from randomSeed 1234567
    (src/params.h:64) -- removals, compaction and reinjected positions bit for bit."""
    distance = 1e-3 * TWO_PI
    rx, ry, rg, rev = reference.dipole_process("periodic", TWO_PI, 1.0, distance, 1, 1234567,
                                               c2.x, c2.y, c2.circulation)
    k = c2_kernel()
    k.uploadDeviceState(VortexSystem(c2.x, c2.y, c2.circulation))
    k.setDipoleRandom(1234567)
    ev = k.processDipoles(distance, "independent")
    assert ev == rev and ev > 0
    out = k.downloadDeviceSystem()
    assert out.size() == len(rx) == c2.n
    assert np.array_equal(out.x, rx) and np.array_equal(out.y, ry)
    assert np.array_equal(out.circulation, rg)


@pytest.mark.timeout(900)
def test_c2_dopri5_step_matches_the_oracle(oracle, c2):
    """One DOPRI5 step of the configuration (rtol 1e-9, atol 1e-12, the default step bounds)
    from dt = 1e-7, well inside the accepted range (the configuration's accepted steps are
    about 2e-7): seven full N=16384 evaluations on each side.  The same acceptance, the error
    norm to 1e-5 relative, and the displacement within 1e-11 normwise."""
    x0, y0, g = c2.x, c2.y, c2.circulation
    k = c2_kernel()
    k.uploadDeviceState(VortexSystem(x0, y0, g))
    rg = k.deviceDopri5Step(1e-7, 1e-12, 1e-9, 1e-12, 0.1)
    x, y = x0.copy(), y0.copy()
    ro = oracle.integrator(c2.n).dopri5_step(C2_SPEC, x, y, g, 1e-7, 1e-12, 1e-9, 1e-12, 0.1)
    assert rg.rejectedSteps == ro[3] == 0
    assert rg.acceptedTimeStep == ro[0] == 1e-7
    assert abs(rg.normalizedError - ro[2]) <= 1e-5 * ro[2]
    assert abs(rg.suggestedTimeStep - ro[1]) <= 1e-5 * ro[1]
    out = k.downloadDeviceSystem()
    assert normwise(out.x - x0, out.y - y0, x - x0, y - y0) <= 1e-11


# ------------------------------------------------------------------ the 4096-instance ensemble


@pytest.fixture(scope="module")
def c4():
    pops = wl.ensemble(4096)
    cfgs = []
    for p in pops:
        if p.geometry == "periodic":
            cfgs.append(InstanceConfig("periodic", box_length=p.box_length))
        elif p.geometry == "disk":
            cfgs.append(InstanceConfig("disk", disk_radius=p.radius))
        else:
            cfgs.append(InstanceConfig("infinite"))
    return cfgs, pops


def c4_ensemble(c4, precision="fp64"):
    cfgs, pops = c4
    e = Ensemble(sum(p.n for p in pops), len(pops), precision)
    e.load(cfgs, [p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
    return e


def ospec(c):
    return spec(c.geometry, core_radius=c.core_radius, box_length=c.box_length,
                image_layers=c.image_layers, radius=c.disk_radius)


def stratified(pops, counts=(("infinite", 32), ("periodic", 16), ("disk", 16))):
    """Per geometry, `count` instances spread evenly over the N-sorted instances."""
    out = []
    for geo, count in counts:
        idx = sorted((k for k, p in enumerate(pops) if p.geometry == geo),
                     key=lambda k: (pops[k].n, k))
        pick = np.linspace(0, len(idx) - 1, count).round().astype(int)
        out += [idx[q] for q in pick]
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c4_evaluate_on_a_stratified_sample(oracle, c4, precision):
    cfgs, pops = c4
    us, vs = c4_ensemble(c4, precision).evaluate()
    sample = stratified(pops)
    assert len(set(sample)) == 64
    for k in sample:
        c, p = cfgs[k], pops[k]
        ur, vr = oracle.evaluate(ospec(c), p.x, p.y, p.circulation)
        if precision == "mixed":
            gate = 1e-6
        else:
            gate = fp64_gate(c.geometry, p.x, p.y, p.circulation, ur, vr)
            if c.geometry == "disk":
                ue, ve = disk_truth(p.x, p.y, p.circulation)
                assert normwise(us[k], vs[k], ue, ve) <= 1e-12, (k, p.n)
        assert normwise(us[k], vs[k], ur, vr) <= gate, (k, c.geometry, p.n)


def oracle_failing_step(oracle, c, x, y, g, dt, steps):
    """RK4 steps of the oracle until its state fails; returns (steps completed, message)."""
    x, y = x.copy(), y.copy()
    it = oracle.integrator(len(x))
    for s in range(steps):
        try:
            it.rk4_step(ospec(c), x, y, g, dt)
        except Exception as exc:      # the reference throws, the run stops
            return s, str(exc)
    return steps, ""


@pytest.mark.timeout(900)
def test_c4_failing_disk_instances_fail_in_the_oracle_too(oracle, c4):
    """The 500-step run of the ensemble (dt = 1e-3, two-step graph replays): instances that
    stop are disk instances whose vortex near the wall is carried through it.  For the eight
    smallest failing instances the oracle's RK4 from the same initial state must also throw
    "vortex lies on or outside the disk", and the ensemble's failing step must lie within the
    window of failing steps of the oracle and its four one-ulp twins (the first vortex's x and
    y moved one ulp either way), which measures how far rounding can move that step."""
    cfgs, pops = c4
    e = c4_ensemble(c4)
    e.rk4(1e-3, 500, use_graph=True)
    status = e.instance_status()
    failing = [k for k, (f, _) in enumerate(status) if f]
    assert failing and all(pops[k].geometry == "disk" for k in failing)
    assert all(status[k][0] == "vortex lies on or outside the disk" for k in failing)
    for k in sorted(failing, key=lambda q: (pops[q].n, q))[:8]:
        c, p = cfgs[k], pops[k]
        done = status[k][1]
        steps, msg = oracle_failing_step(oracle, c, p.x, p.y, p.circulation, 1e-3, 500)
        assert msg == "vortex lies on or outside the disk", (k, p.n, steps, msg)
        window = [steps]
        for coord, sign in (("x", 1), ("x", -1), ("y", 1), ("y", -1)):
            x, y = p.x.copy(), p.y.copy()
            a = x if coord == "x" else y
            a[0] = np.nextafter(a[0], sign * np.inf)
            s_twin, m_twin = oracle_failing_step(oracle, c, x, y, p.circulation, 1e-3, 500)
            window.append(s_twin if m_twin else 500)
        assert min(window) <= done <= max(window), (k, p.n, done, window)


# ------------------------------------------------------------------ serving with large runs


@pytest.mark.timeout(1200)
def test_c5_serving_admits_large_runs_under_a_binding_budget(oracle):
    """The serving loop with runs at N=131072 and N=65536 (disk populations of the benchmark's
    kind, FP64 at rtol 1e-8) beside small runs of both precisions, under a budget that cannot
    hold both large runs at once (the larger run's footprint plus half the smaller one's, about
    1.2-2.4 GB by the tile the launch rules give them): the second waits until the first has
    left.  The
    device allocation never exceeds the budget; each large run gives the same bits as when
    served alone; and the final state's FP64 velocities agree with the oracle on the first 256
    targets and the wall shell (the last 256)."""
    from pointvortex_amd.serve import Request, Server

    # From dt = 1e-6 the controller needs at least six accepted steps (at most 5x growth per
    # step) to reach t = 1e-3, more where the closest pairs limit the step.
    big = [Request(0, 0.0, "disk", 131072, 71, 1e-3, 1e-8, 1e-11, dt0=1e-6),
           Request(1, 0.0, "disk", 65536, 72, 1e-3, 1e-8, 1e-11, dt0=1e-6)]
    small = [Request(2 + k, 0.0, ["infinite", "periodic", "disk"][k % 3], 64 + 96 * k,
                     80 + k, 2e-4, 1e-8 if k % 2 else 1e-5, 1e-11 if k % 2 else 1e-8)
             for k in range(10)]
    reqs = big + small
    probe = Server(hbm_budget_bytes=int(8e9))
    budget = probe.footprint(131072) + probe.footprint(65536) // 2
    srv = Server(hbm_budget_bytes=budget, max_runs=16, keep_results=True)
    assert srv.footprint(131072) + srv.footprint(65536) > budget
    stats = srv.run(reqs, max_wall=600.0)
    res = {r.rid: r for r in stats.results}
    assert len(res) == len(reqs) and all(r.state == "finished" for r in res.values()), \
        [(r.rid, r.state, r.failure) for r in stats.results]
    assert 0 < srv.peak_device_bytes <= budget
    assert res[1].admitted >= res[0].finished          # the budget bound
    assert {r.precision for r in res.values()} == {"fp64", "mixed"}
    for req in big:
        r = res[req.rid]
        assert 6 <= r.accepted + r.rejected <= 600, (req.n, r.accepted, r.rejected)
        alone = Server(hbm_budget_bytes=budget, max_runs=2, keep_results=True)
        a = alone.run([req], max_wall=600.0).results[0]
        assert a.state == "finished"
        assert (a.accepted, a.rejected) == (r.accepted, r.rejected)
        assert np.array_equal(a.x, r.x) and np.array_equal(a.y, r.y), req.n
        pop, _ = req.population()
        assert np.all(r.x * r.x + r.y * r.y < 1.0)
        k = HipKernel(SimParams(boundaryCondition="disk"))
        vel = k.evaluate(VortexSystem(r.x, r.y, pop.circulation))
        n = req.n
        for b, e in [(0, 256), (n - 256, n)]:
            ur, vr = oracle.evaluate_range(spec("disk"), r.x, r.y, pop.circulation, b, e)
            from tests.conftest import disk_longdouble_targets
            ue, ve = disk_longdouble_targets(r.x, r.y, pop.circulation, np.arange(b, e))
            assert normwise(vel.x[b:e], vel.y[b:e], ue, ve) <= 1e-12, (n, b)
            gate = max(1e-12, 2.0 * normwise(ur[b:e], vr[b:e], ue, ve))
            assert normwise(vel.x[b:e], vel.y[b:e], ur[b:e], vr[b:e]) <= gate, (n, b)
