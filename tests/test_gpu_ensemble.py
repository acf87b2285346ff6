"""Ensembles: every instance against the CPU oracle, hipGraph replay against eager launches,
and adaptive runs alone against the same runs in a mix (they must agree bit for bit: each
target's sum is computed by one wave in a fixed order whatever else is scheduled)."""
import numpy as np
import pytest

from oracle.pyoracle import spec
from pointvortex_amd import HipKernel, InvalidArgument, SimParams, VortexSystem
from pointvortex_amd import workloads as wl
from pointvortex_amd.ensemble import Ensemble, InstanceConfig, bytes_per_vortex
from tests.conftest import fp64_gate, normwise

pytestmark = pytest.mark.gpu


def mixed_population(count, seed, n_min=2, n_max=300):
    rng = np.random.default_rng(seed)
    cfgs, pops = [], []
    for k in range(count):
        geo = ["infinite", "periodic", "disk"][k % 3]
        n = int(rng.integers(n_min, n_max + 1))
        n += n % 2
        if geo == "infinite":
            p = wl.infinite(n, 100 + k)
            cfgs.append(InstanceConfig("infinite", core_radius=0.01 if k % 2 else 0.0))
        elif geo == "periodic":
            p = wl.periodic(n, 100 + k, box_length=1.0)
            cfgs.append(InstanceConfig("periodic", box_length=1.0, image_layers=8))
        else:
            p = wl.disk(n, 100 + k, wall_fraction=0.0)
            cfgs.append(InstanceConfig("disk", disk_radius=1.0))
        pops.append(p)
    return cfgs, pops


def ospec(c):
    return spec(c.geometry, core_radius=c.core_radius, box_length=c.box_length,
                image_layers=c.image_layers, radius=c.disk_radius)


def oracle_rk4(oracle, c, x, y, g, dt, steps):
    x, y = x.copy(), y.copy()
    it = oracle.integrator(len(x))
    for _ in range(steps):
        it.rk4_step(ospec(c), x, y, g, dt)
    return x, y


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-6)])
def test_ensemble_evaluate_matches_oracle(oracle, precision, tol, mode):
    cfgs, pops = mixed_population(30, 1)
    e = Ensemble(sum(p.n for p in pops), 64, precision)
    e.set_symmetric(mode)
    e.load(cfgs, [p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
    us, vs = e.evaluate()
    for c, p, u, v in zip(cfgs, pops, us, vs):
        ur, vr = oracle.evaluate(ospec(c), p.x, p.y, p.circulation)
        gate = fp64_gate(c.geometry, p.x, p.y, p.circulation, ur, vr) if tol < 1e-6 else tol
        assert normwise(u, v, ur, vr) <= gate, (c, p.n)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-6)])
def test_ensemble_periodic_layers_per_instance(oracle, precision, tol, mode):
    """Each instance carries its own image-layer count, and with it its own closed-form and
    series weights: L = 0, 1, 2, 3, 8, 20 side by side, in two box lengths."""
    cfgs, pops = [], []
    for k, layers in enumerate([0, 1, 2, 3, 8, 20] * 2):
        box = 1.0 if k < 6 else 2.5
        pops.append(wl.periodic(60 + 20 * k, 300 + k, box_length=box))
        cfgs.append(InstanceConfig("periodic", box_length=box, image_layers=layers))
    e = Ensemble(sum(p.n for p in pops), 16, precision)
    e.set_symmetric(mode)
    e.load(cfgs, [p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
    us, vs = e.evaluate()
    for c, p, u, v in zip(cfgs, pops, us, vs):
        ur, vr = oracle.evaluate(ospec(c), p.x, p.y, p.circulation)
        assert normwise(u, v, ur, vr) <= tol, (c, p.n)


def test_ensemble_rk4_graph_matches_eager_and_oracle(oracle):
    cfgs, pops = mixed_population(24, 2)
    args = ([p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
    total = sum(p.n for p in pops)
    a = Ensemble(total, 32)
    b = Ensemble(total, 32)
    a.load(cfgs, *args)
    b.load(cfgs, *args)
    a.rk4(1e-3, 6, use_graph=True)
    b.rk4(1e-3, 6, use_graph=False)
    xa, ya = a.read()
    xb, yb = b.read()
    for k in range(len(pops)):
        assert np.array_equal(xa[k], xb[k]) and np.array_equal(ya[k], yb[k])
    for k, (c, p, x, y) in enumerate(zip(cfgs, pops, xa, ya)):
        xr, yr = oracle_rk4(oracle, c, p.x, p.y, p.circulation, 1e-3, 6)
        # The FP64 velocity bar (normwise 1e-12, or the oracle's own distance from an
        # extended-precision sum near a disk wall) carries over to the displacement as far as
        # the instance's conditioning allows: a close pair spinning at omega * dt >> 1 makes
        # RK4 amplify any velocity difference.  The oracle measures that amplification itself
        # by rerunning with circulations perturbed by the bar (velocity is linear in them).
        ur, vr = oracle.evaluate(ospec(c), p.x, p.y, p.circulation)
        gate = fp64_gate(c.geometry, p.x, p.y, p.circulation, ur, vr)
        rng = np.random.default_rng(k)
        g = p.circulation * (1.0 + gate * rng.standard_normal(p.n))
        if c.geometry == "periodic":
            g -= g.mean()
        xq, yq = oracle_rk4(oracle, c, p.x, p.y, g, 1e-3, 6)
        amplified = normwise(xq - p.x, yq - p.y, xr - p.x, yr - p.y)
        err = normwise(x - p.x, y - p.y, xr - p.x, yr - p.y)
        assert err <= max(100 * gate, 10 * amplified), (k, c, p.n, err, gate, amplified)


def test_ensemble_graph_replays_repeat(oracle):
    """Replays of one captured graph, with eager work and synchronisation in between, give the
    same states as eager launches."""
    cfgs, pops = mixed_population(9, 3)
    args = ([p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
    total = sum(p.n for p in pops)
    a, b = Ensemble(total, 16), Ensemble(total, 16)
    a.load(cfgs, *args)
    b.load(cfgs, *args)
    for _ in range(3):
        a.rk4(1e-3, 4, use_graph=True)
        a.evaluate()
        b.rk4(1e-3, 4, use_graph=False)
        b.evaluate()
        xa, _ = a.read()
        xb, _ = b.read()
        assert all(np.array_equal(p, q) for p, q in zip(xa, xb))


@pytest.mark.parametrize("counts", [(3, 2), (1, 2, 2), (3, 3, 4)])
def test_ensemble_graph_after_an_odd_step_count(counts):
    """An odd step count leaves the state in the other ping-pong buffer (its last step runs
    eagerly); the next graph call with the same dt must step from there, not replay the graph
    captured for the first orientation.  Every sequence equals the same number of eager
    steps, bit for bit, and the step counter agrees."""
    cfgs, pops = mixed_population(9, 5)
    args = ([p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
    total = sum(p.n for p in pops)
    a, b = Ensemble(total, 16), Ensemble(total, 16)
    a.load(cfgs, *args)
    b.load(cfgs, *args)
    for c in counts:
        a.rk4(1e-3, c, use_graph=True)
    b.rk4(1e-3, sum(counts), use_graph=False)
    xa, ya = a.read()
    xb, yb = b.read()
    for k in range(len(pops)):
        assert np.array_equal(xa[k], xb[k]) and np.array_equal(ya[k], yb[k]), k
    assert all(d == sum(counts) for f, d in a.instance_status() if not f)


def drive(ens, max_attempts):
    """Attempts until no run is left running; a bounded loop, so a run that stops making
    progress fails the test instead of spinning."""
    for _ in range(max_attempts):
        if not any(s.state == "running" for s in ens.poll()):
            return
        ens.attempt()
    raise AssertionError(f"runs still running after {max_attempts} attempts: {ens.poll()}")


def run_alone_and_mixed(precision, mode=-1):
    rng = np.random.default_rng(7)
    target = wl.periodic(300 if mode == 1 else 64, 11, box_length=1.0)
    if mode == 1:
        # Three tiles of the symmetric decomposition; weaker circulations keep 300 vortices'
        # closest pairs as slow as the 64-vortex run's.
        target.circulation *= 0.01
    cfg = InstanceConfig("periodic", box_length=1.0)
    kw = dict(t_end=0.002, dt0=1e-4, atol=1e-10, rtol=1e-8)
    alone = Ensemble(4096, 8, precision)
    alone.set_symmetric(mode)
    s0 = alone.admit(cfg, target.x, target.y, target.circulation, **kw)
    drive(alone, 2000)
    st_alone = [s for s in alone.poll() if s.slot == s0][0]
    xa, ya = alone.fetch(s0, target.n)
    mix = Ensemble(20000, 16, precision)
    mix.set_symmetric(mode)
    others = []
    for k in range(6):
        p = wl.disk(int(rng.integers(50, 900)), 50 + k, wall_fraction=0.0, gaussian=True)
        others.append(mix.admit(InstanceConfig("disk"), p.x, p.y, p.circulation, t_end=0.01,
                                rtol=1e-7))
    mix.attempt()
    mix.attempt()
    s1 = mix.admit(cfg, target.x, target.y, target.circulation, **kw)
    drive(mix, 2000)
    st_mix = [s for s in mix.poll() if s.slot == s1][0]
    xm, ym = mix.fetch(s1, target.n)
    return st_alone, st_mix, (xa, ya), (xm, ym), mix.poll()


@pytest.mark.parametrize("mode", [-1, 1])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_adaptive_run_alone_equals_run_in_the_mix(precision, mode):
    st_alone, st_mix, (xa, ya), (xm, ym), everyone = run_alone_and_mixed(precision, mode)
    assert st_alone.state == st_mix.state == "finished"
    assert st_alone.accepted == st_mix.accepted and st_alone.rejected == st_mix.rejected
    assert np.array_equal(xa, xm) and np.array_equal(ya, ym)
    assert all(s.state == "finished" for s in everyone)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_attempts_match_single_attempts(precision):
    """pv_batch_attempts(b, k) against k single attempts: the same states and positions bit for
    bit, including runs that finish part-way through a batch (they sit out the rest)."""
    cfgs, pops = mixed_population(5, 21, n_min=20, n_max=200)
    ends = [0.002, 0.0003, 0.002, 0.0005, 0.002]

    def make():
        e = Ensemble(4096, 8, precision)
        e.set_symmetric(1)
        return e, [e.admit(c, p.x, p.y, p.circulation, t_end=t, dt0=1e-4, atol=1e-10, rtol=1e-8)
                   for c, p, t in zip(cfgs, pops, ends)]

    one, s1 = make()
    many, s8 = make()
    for _ in range(24):
        one.attempt()
    for _ in range(3):
        many.attempt(8)
    p1 = {s.slot: s for s in one.poll()}
    p8 = {s.slot: s for s in many.poll()}
    for a, b, p in zip(s1, s8, pops):
        assert (p1[a].state, p1[a].t, p1[a].dt, p1[a].accepted, p1[a].rejected) == \
            (p8[b].state, p8[b].t, p8[b].dt, p8[b].accepted, p8[b].rejected)
        xa, ya = one.fetch(a, p.n)
        xb, yb = many.fetch(b, p.n)
        assert np.array_equal(xa, xb) and np.array_equal(ya, yb)
    assert sum(p1[a].state == "finished" for a in s1) >= 1
    assert sum(p1[a].state == "running" for a in s1) >= 1
    with pytest.raises(InvalidArgument, match="attempt count must be positive"):
        many.attempt(0)


def test_adaptive_matches_single_handle_dopri5():
    """One adaptive run against the single-handle DOPRI5 path driven to the same end time with
    the reference's landing rule (src/main.cpp:312-314)."""
    p = wl.infinite(100, 5)
    kw = dict(atol=1e-12, rtol=1e-9, min_dt=1e-12, max_dt=0.1)
    e = Ensemble(1000, 4)
    slot = e.admit(InstanceConfig("infinite"), p.x, p.y, p.circulation, t_end=0.002, dt0=1e-4,
                   **kw)
    drive(e, 500)
    xe, ye = e.fetch(slot, p.n)
    k = HipKernel(SimParams(boundaryCondition="infinite"))
    s = VortexSystem(p.x, p.y, p.circulation)
    k.uploadDeviceState(s)
    t, dt, steps = 0.0, 1e-4, 0
    for _ in range(500):
        if t >= 0.002:
            break
        r = k.deviceDopri5Step(min(dt, 0.002 - t), kw["atol"], kw["rtol"], kw["min_dt"],
                               kw["max_dt"])
        t += r.acceptedTimeStep
        dt = r.suggestedTimeStep
        steps += 1
        if t + 16 * 2.220446049250313e-16 * abs(t) >= 0.002:
            break
    out = s.copy()
    k.downloadDeviceState(out)
    st = [q for q in e.poll() if q.slot == slot][0]
    assert st.accepted == steps
    scale = np.max(np.abs(np.concatenate([out.x, out.y])))
    assert np.max(np.abs(xe - out.x)) <= 1e-10 * scale
    assert np.max(np.abs(ye - out.y)) <= 1e-10 * scale


def test_admission_by_footprint_and_release():
    e = Ensemble(1000, 4)
    p = wl.infinite(600, 1)
    q = wl.infinite(500, 2)
    a = e.admit(InstanceConfig(), p.x, p.y, p.circulation, t_end=1e-3)
    assert a is not None
    assert e.admit(InstanceConfig(), q.x, q.y, q.circulation, t_end=1e-3) is None
    e.release(a)
    b = e.admit(InstanceConfig(), q.x, q.y, q.circulation, t_end=1e-3)
    assert b is not None
    assert bytes_per_vortex() >= 200


def test_failed_run_does_not_stop_others():
    e = Ensemble(2000, 4)
    good = wl.infinite(100, 3)
    a = e.admit(InstanceConfig(), good.x, good.y, good.circulation, t_end=0.005)
    bad_x = good.x.copy()
    bad_x[1] = bad_x[0]
    bad_y = good.y.copy()
    bad_y[1] = bad_y[0]
    b = e.admit(InstanceConfig(), bad_x, bad_y, good.circulation, t_end=0.005)
    drive(e, 2000)
    st = {s.slot: s for s in e.poll()}
    assert st[a].state == "finished"
    assert st[b].state == "failed"
    assert st[b].failure == "coincident vortices in singular point-vortex kernel"


COINCIDENT = {"infinite": "coincident vortices in singular point-vortex kernel",
              "periodic": "coincident periodic vortices",
              "disk": "coincident vortices in disk"}


@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("mode", [0, 1])
def test_coincident_pair_in_one_instance_names_its_geometry(mode, adaptive):
    """A coincident pair in one instance fails that instance with the reference's message for
    its geometry (src/compute.cpp:131-132, 254-255, 334-335), as the single-system path
    reports it; an overflowing velocity without a zero denominator keeps the non-finite
    message; the other instances are unaffected."""
    pops = {"infinite": wl.infinite(80, 1), "periodic": wl.periodic(80, 2, box_length=1.0),
            "disk": wl.disk(80, 3, wall_fraction=0.0)}
    cfgs = {"infinite": InstanceConfig("infinite"),
            "periodic": InstanceConfig("periodic", box_length=1.0),
            "disk": InstanceConfig("disk")}
    members = []
    for geo in ("infinite", "periodic", "disk"):
        p = pops[geo]
        members.append((cfgs[geo], p.x, p.y, p.circulation))
        x, y = p.x.copy(), p.y.copy()
        x[5], y[5] = x[11], y[11]
        members.append((cfgs[geo], x, y, p.circulation))
    # Circulations near the largest double: the sum overflows, no denominator is zero.
    p = wl.infinite(40, 4)
    members.append((cfgs["infinite"], p.x, p.y, p.circulation * 1e308))
    e = Ensemble(sum(len(m[1]) for m in members), 16)
    e.set_symmetric(mode)
    want = ["", COINCIDENT["infinite"], "", COINCIDENT["periodic"], "", COINCIDENT["disk"],
            "non-finite velocity; check scales and close encounters"]
    if not adaptive:
        e.load(*[list(t) for t in zip(*members)])
        e.rk4(1e-4, 2, use_graph=False)
        assert e.failures() == want
        return
    slots = [e.admit(c, x, y, g, t_end=2e-4, dt0=1e-4) for c, x, y, g in members]
    drive(e, 200)
    st = {s.slot: s for s in e.poll()}
    assert [st[k].failure for k in slots] == want
    assert [st[k].state for k in slots] == ["finished" if w == "" else "failed" for w in want]


def test_serving_loop_completes_and_matches_alone():
    """Requests arriving while others run: every result equals the same request served alone."""
    from pointvortex_amd.serve import Server, make_requests

    reqs = make_requests(12, 5, rate=400.0, n_min=32, n_max=200, t_min=0.0005, t_max=0.002,
                         rtol_min=1e-8, rtol_max=1e-6)
    srv = Server(hbm_budget_bytes=4_000_000, max_runs=8, keep_results=True)
    stats = srv.run(reqs, max_wall=60.0)
    summary = stats.summary()
    assert summary["completed"] == len(reqs) and summary["failed"] == 0
    assert summary["p99_s"] >= summary["p50_s"] > 0
    by_id = {r.rid: r for r in stats.results}
    for req in reqs[:4]:
        alone = Server(hbm_budget_bytes=4_000_000, max_runs=2, keep_results=True)
        req0 = type(req)(**{**req.__dict__, "arrival": 0.0})
        r = alone.run([req0], max_wall=30.0).results[0]
        assert r.accepted == by_id[req.rid].accepted
        assert np.array_equal(r.x, by_id[req.rid].x) and np.array_equal(r.y, by_id[req.rid].y)


@pytest.mark.parametrize("mode", [0, 1])
def test_fixed_step_failures_are_per_instance(oracle, mode):
    """A fixed-step instance that fails stops at the input state of its failing step and the
    others go on unchanged: a coincident pair fails in the first step, a dipole driven into
    the disk wall in the seventh.  Nine steps run as four two-step graph replays plus one
    eager step, so both stopped states end up in the buffer the last step swapped away."""
    cfgs, pops = mixed_population(6, 4)
    healthy = [(c, p.x, p.y, p.circulation) for c, p in zip(cfgs, pops)]
    good = wl.infinite(40, 9)
    bx, by = good.x.copy(), good.y.copy()
    bx[3], by[3] = bx[7], by[7]
    coincident = (InstanceConfig("infinite"), bx, by, good.circulation)
    wx, wy = np.array([0.8, 0.8, -0.3, -0.2]), np.array([0.005, -0.005, 0.1, -0.4])
    wg = np.array([1.0, -1.0, 0.5, -0.5])
    wall = (InstanceConfig("disk"), wx, wy, wg)
    members = healthy[:3] + [coincident] + healthy[3:] + [wall]
    e = Ensemble(sum(len(m[1]) for m in members), 16)
    e.set_symmetric(mode)
    e.load(*[list(t) for t in zip(*members)])
    assert e.symmetric == bool(mode)
    ref = Ensemble(sum(len(m[1]) for m in healthy), 16)
    ref.set_symmetric(mode)
    ref.load(*[list(t) for t in zip(*healthy)])
    dt = 2e-3
    e.rk4(dt, 9, use_graph=True)
    ref.rk4(dt, 9, use_graph=True)
    status = e.instance_status()
    fails = [f for f, _ in status]
    assert [d for _, d in status] == [9, 9, 9, 0, 9, 9, 9, 6], status
    assert fails[3] == "coincident vortices in singular point-vortex kernel", fails
    assert fails[7] == "vortex lies on or outside the disk", fails
    assert all(f == "" for k, f in enumerate(fails) if k not in (3, 7)), fails
    xs, ys = e.read()
    xr, yr = ref.read()
    for k_e, k_r in zip([0, 1, 2, 4, 5, 6], range(6)):
        assert np.array_equal(xs[k_e], xr[k_r]) and np.array_equal(ys[k_e], yr[k_r])
    assert np.array_equal(xs[3], bx) and np.array_equal(ys[3], by)
    # The wall instance's state after six steps, against the oracle; the tolerance is the
    # oracle's own response to a 1e-12 relative velocity change over those steps.
    spec_disk = ospec(wall[0])
    ox, oy = oracle_rk4(oracle, wall[0], wx, wy, wg, dt, 6)
    px, py = oracle_rk4(oracle, wall[0], wx, wy, wg * (1 + 1e-12 * np.array([1, -1, 1, -1])),
                        dt, 6)
    amplified = normwise(px - wx, py - wy, ox - wx, oy - wy)
    err = normwise(xs[7] - wx, ys[7] - wy, ox - wx, oy - wy)
    assert err <= max(1e-10, 10 * amplified), (err, amplified)
    with pytest.raises(Exception, match="outside the disk"):
        oracle.integrator(4).rk4_step(spec_disk, ox.copy(), oy.copy(), wg, dt)


@pytest.mark.parametrize("mode", [0, 1])
def test_ensemble_disk_near_centre_source(oracle, mode):
    """The ensemble kernels' per-wave tiles take the reference's image form when a source sits
    within 2^-100 R of the centre (symmetric mode: in tile pairs where that vortex is a target
    and where it is a source)."""
    pops = [wl.disk(150, 61, wall_fraction=0.0), wl.disk(90, 62, wall_fraction=0.0),
            wl.disk(400, 63, wall_fraction=0.0)]
    pops[0].x[17], pops[0].y[17] = 1e-155, 0.0
    pops[1].x[80], pops[1].y[80] = 0.0, 0.0
    pops[2].x[200], pops[2].y[200] = -2e-160, 1e-161
    cfgs = [InstanceConfig("disk")] * 3
    e = Ensemble(sum(p.n for p in pops), 4)
    e.set_symmetric(mode)
    e.load(cfgs, [p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
    us, vs = e.evaluate()
    for c, p, u, v in zip(cfgs, pops, us, vs):
        ur, vr = oracle.evaluate(ospec(c), p.x, p.y, p.circulation)
        assert normwise(u, v, ur, vr) <= fp64_gate("disk", p.x, p.y, p.circulation, ur, vr)


@pytest.mark.parametrize("pos", [(2e-160, 1e-161), (-3e-31, 2e-31), (6e-13, -4e-13),
                                 (8e-13, 5e-13), (3e-7, 1e-7)])
def test_ensemble_mixed_disk_near_centre_source(oracle, pos):
    """Mixed disk items share one FP32 reciprocal between the two image terms (w = 1/|x|^2
    staged with each source) unless the item holds a vortex within 2^-40 R of the centre,
    which takes the direct form.  One such vortex per instance, in a middle tile (a source
    for earlier tiles and a target for later ones), on both sides of that bound, at 64-, 128-
    and 256-vortex tiles; every other instance of the launch is away from the centre."""
    pops = [wl.disk(n, 74 + k, wall_fraction=0.0) for k, n in enumerate((700, 1100, 1800))]
    for p in pops:
        p.x[p.n // 2], p.y[p.n // 2] = pos
    cfgs = [InstanceConfig("disk")] * 3
    for tpl in (1, 2, 4):
        e = Ensemble(sum(p.n for p in pops), 4, "mixed")
        e.set_launch_table([("disk", "mixed", 2**64 - 1, tpl, 4)])
        e.load(cfgs, [p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
        assert e.symmetric
        us, vs = e.evaluate()
        e.close()
        for c, p, u, v in zip(cfgs, pops, us, vs):
            assert np.all(np.isfinite(u)) and np.all(np.isfinite(v))
            ur, vr = oracle.evaluate(ospec(c), p.x, p.y, p.circulation)
            assert normwise(u, v, ur, vr) <= 1e-6, (tpl, p.n)
            i = p.n // 2
            assert np.hypot(u[i] - ur[i], v[i] - vr[i]) <= 1e-5 * np.hypot(ur[i], vr[i])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_symmetric_and_ordered_ensembles_agree(precision):
    """The two work decompositions of one ensemble (instances up to 5 tiles of 128): FP64
    velocities within the FP64 bar of each other, mixed within the mixed bar of FP64;
    symmetric repeats are bit-identical, and switching the mode keeps the loaded state."""
    cfgs, pops = mixed_population(18, 21, n_min=2, n_max=620)
    args = ([p.x for p in pops], [p.y for p in pops], [p.circulation for p in pops])
    total = sum(p.n for p in pops)
    e = Ensemble(total, 32, precision)
    e.load(cfgs, *args)
    assert e.symmetric                      # fixed-step default
    us, vs = e.evaluate()
    us2, vs2 = e.evaluate()
    assert all(np.array_equal(a, b) for a, b in zip(us, us2))
    assert all(np.array_equal(a, b) for a, b in zip(vs, vs2))
    e.set_symmetric(0)
    assert not e.symmetric
    uo, vo = e.evaluate()
    ref = Ensemble(total, 32, "fp64")
    ref.set_symmetric(0)
    ref.load(cfgs, *args)
    ur, vr = ref.evaluate()
    bar = 1e-12 if precision == "fp64" else 1e-6
    for k, (c, p) in enumerate(zip(cfgs, pops)):
        assert normwise(us[k], vs[k], ur[k], vr[k]) <= bar, (k, c, p.n)
        assert normwise(uo[k], vo[k], ur[k], vr[k]) <= bar, (k, c, p.n)
    e.set_symmetric(1)
    e.rk4(1e-4, 2, use_graph=True)
    ref.set_symmetric(1)
    ref.rk4(1e-4, 2, use_graph=False)
    if precision == "fp64":
        xa, ya = e.read()
        xb, yb = ref.read()
        assert all(np.array_equal(a, b) for a, b in zip(xa, xb))
        assert all(np.array_equal(a, b) for a, b in zip(ya, yb))


def test_serving_budget_counts_both_precisions_and_partials():
    """The footprint of every admitted run (state plus symmetric partials, FP64 and mixed
    engines together) never exceeds the budget; runs wait for room and all complete."""
    from pointvortex_amd.serve import Server, make_requests

    reqs = make_requests(6, 9, rate=0.0, n_min=500, n_max=700, t_min=2e-4, t_max=4e-4,
                         rtol_min=1e-8, rtol_max=1e-5)
    probe = Server(hbm_budget_bytes=1 << 30)
    big = max(probe.request_footprint(r) for r in reqs)
    assert big > 700 * bytes_per_vortex()          # the partials are counted
    srv = Server(hbm_budget_bytes=int(2.5 * big), max_runs=8)
    stats = srv.run(reqs, max_wall=60.0)
    assert stats.summary()["completed"] == len(reqs)
    assert srv.peak_used <= srv.budget
    assert srv.peak_used >= 2 * min(srv.request_footprint(r) for r in reqs)   # overlapped
    assert srv.used == 0
    # What the one ensemble actually held on the device, both precisions in it at once.
    assert {r.precision for r in stats.results} == {"fp64", "mixed"}
    assert srv.mixed_and_fp64_at_once
    assert 0 < srv.peak_device_bytes <= srv.budget


def test_mixed_and_fp64_runs_in_one_ensemble_equal_their_runs_alone():
    """FP64 and mixed runs admitted into one ensemble advance in the same launches (one kernel
    dispatching on each run's precision), and each gives the same bits as alone in an
    ensemble of its own precision."""
    cfgs, pops = mixed_population(6, 31, n_min=40, n_max=300)
    precs = ["fp64", "mixed"] * 3
    kw = dict(t_end=0.002, dt0=1e-4, atol=1e-10, rtol=1e-8)
    both = Ensemble(4096, 8)
    both.set_symmetric(1)
    slots = [both.admit(c, p.x, p.y, p.circulation, precision=q, **kw)
             for c, p, q in zip(cfgs, pops, precs)]
    drive(both, 3000)
    st = {s.slot: s for s in both.poll()}
    for c, p, q, k in zip(cfgs, pops, precs, slots):
        alone = Ensemble(1024, 2, q)
        alone.set_symmetric(1)
        a = alone.admit(c, p.x, p.y, p.circulation, **kw)
        drive(alone, 3000)
        sa = alone.poll()[0]
        assert sa.state == st[k].state == "finished"
        assert (sa.accepted, sa.rejected) == (st[k].accepted, st[k].rejected)
        xa, ya = alone.fetch(a, p.n)
        xb, yb = both.fetch(k, p.n)
        assert np.array_equal(xa, xb) and np.array_equal(ya, yb), (c, p.n, q)


def test_byte_limit_bounds_what_the_ensemble_holds():
    """With a byte limit, admission refuses a run whose state, items and partial sums would
    take the ensemble's device allocation past it; after the others leave it fits.  (Runs of
    about 20000 vortices, whose partial sums outweigh their state.)"""
    e = Ensemble(100_000, 8)
    e.set_symmetric(1)
    p, q = wl.infinite(20000, 1), wl.infinite(18000, 2)
    e.set_byte_limit(0)
    a = e.admit(InstanceConfig(), p.x, p.y, p.circulation, t_end=1e-4)
    e.attempt()
    one = e.device_bytes()
    e.release(a)
    e.set_byte_limit(int(1.5 * one))
    a = e.admit(InstanceConfig(), p.x, p.y, p.circulation, t_end=1e-4)
    assert a is not None
    assert e.admit(InstanceConfig(), q.x, q.y, q.circulation, t_end=1e-4) is None
    e.attempt()
    assert e.device_bytes() <= int(1.5 * one)
    e.release(a)
    b = e.admit(InstanceConfig(), q.x, q.y, q.circulation, t_end=1e-4)
    assert b is not None
    e.attempt()
    assert e.device_bytes() <= int(1.5 * one)
