"""Parity of the HIP backend (through the C-ABI) against the CPU FP64 oracle.

Tolerances:
* FP64 velocities: normwise relative error <= 1e-12 (the reference's own FP64 error against
  an exact sum is ~1e-14 normwise); per-vortex relative error (floor 1e-3 rms) <= 1e-10.
* mixed velocities: normwise <= 1e-6 against the FP64 oracle.
* Hamiltonians: within 16 * 2^-53 S of the extended-precision value, S the sum of the terms'
  magnitudes (a bound from the rounding of each term and a compensated sum); moments 1e-13.
"""
import numpy as np
import pytest

from oracle.pyoracle import spec
from pointvortex_amd import (BackendRuntimeError, HipKernel, InvalidArgument, LogicError,
                             OutOfRange, SimParams, VortexSystem)
from pointvortex_amd import workloads as wl
from tests.conftest import disk_truth, fp64_gate, hamiltonian_longdouble, normwise, per_vortex

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * np.pi


def kernel_for(geometry, precision="fp64", **kw):
    p = SimParams(boundaryCondition=geometry, precision=precision)
    for k, v in kw.items():
        setattr(p, k, v)
    return HipKernel(p)


def spec_for(geometry, **kw):
    return spec(geometry, core_radius=kw.get("coreRadius", 0.0),
                box_length=kw.get("boxLengthX", 2.0), image_layers=kw.get("periodicImageLayers", 8),
                radius=kw.get("diskRadius", 1.0))


def population(geometry, n, seed):
    if geometry == "infinite":
        return wl.infinite(n, seed)
    if geometry == "periodic":
        return wl.periodic(n, seed, box_length=2.0)
    return wl.disk(n, seed)


# ------------------------------------------------------------------ known answers

def test_single_and_pair_known_answers():
    """tests/tests.cpp:26-41: single vortex at rest; co-rotating pair v = -+1/(4 pi)."""
    k = kernel_for("infinite")
    vel = k.evaluate(VortexSystem([2.0], [-3.0], [4.0]))
    assert vel.x[0] == 0.0 and vel.y[0] == 0.0
    vel = k.evaluate(VortexSystem([-1.0, 1.0], [0.0, 0.0], [1.0, 1.0]))
    speed = 1.0 / (4.0 * np.pi)
    assert abs(vel.y[0] + speed) <= 1e-15 and abs(vel.y[1] - speed) <= 1e-15


def test_periodic_translation_invariance():
    """tests/tests.cpp:126-140 (L=2, 10 layers)."""
    k = kernel_for("periodic", boxLengthX=2.0, boxLengthY=2.0, periodicImageLayers=10)
    s = VortexSystem([-0.4, 0.4], [0.1, -0.1], [1.0, -1.0])
    a = k.evaluate(s)
    s2 = VortexSystem([-0.4 + 2.0, 0.4], [0.1, -0.1 - 2.0], [1.0, -1.0])
    b = k.evaluate(s2)
    assert np.max(np.abs(a.x - b.x)) <= 2e-14 and np.max(np.abs(a.y - b.y)) <= 2e-14


def test_disk_single_vortex_image():
    """tests/tests.cpp:141-153: r = 0.5, R = 2: u = 0, v = r / (2 pi (R^2 - r^2))."""
    k = kernel_for("disk", diskRadius=2.0)
    vel = k.evaluate(VortexSystem([0.5], [0.0], [1.0]))
    assert abs(vel.x[0]) <= 1e-15
    assert abs(vel.y[0] - 0.5 / (2.0 * np.pi * (4.0 - 0.25))) <= 1e-15


def test_close_encounters():
    """tests/audit_tests.cpp:27-47."""
    inv2pi = 0.5 / np.pi
    k = kernel_for("periodic", boxLengthX=2.0, boxLengthY=2.0)
    pair = VortexSystem([0.0, 1e-10], [0.0, 0.0], [1.0, -1.0])
    vel = k.evaluate(pair)
    assert abs(vel.y[0] / (inv2pi / 1e-10) - 1.0) <= 1e-12
    assert np.isfinite(k.hamiltonian(pair))           # audit_tests.cpp:35-36
    with pytest.raises(BackendRuntimeError, match="coincident periodic vortices"):
        k.evaluate(VortexSystem([0.0, 0.0], [0.0, 0.0], [1.0, -1.0]))
    d = kernel_for("disk", diskRadius=1.0)
    vel = d.evaluate(VortexSystem([0.2, 1e-155], [0.0, 0.0], [1.0, 1.0]))
    assert abs(vel.x[0]) <= 1e-15
    assert abs(vel.y[0] - inv2pi * (5 + 0.2 / 0.96)) <= 1e-14
    assert abs(vel.y[1] - inv2pi * (-5 + 0.2)) <= 1e-14
    with pytest.raises(InvalidArgument):
        kernel_for("periodic", boxLengthX=2.0, boxLengthY=2.0, periodicImageLayers=65)


@pytest.mark.parametrize("symmetric", [0, 1])
@pytest.mark.parametrize("precision,seps,bound", [
    ("fp64", [1e-2, 1e-6, 1e-12, 1e-50, 1e-150], 1e-10),
    ("mixed", [1e-2, 1e-6, 1e-12], 1e-5)])
def test_periodic_close_pairs(oracle, precision, seps, bound, symmetric):
    """A pair at separations down to 1e-150 (FP64) in a population, on both paths.  The rows
    m = 0, +-1 share one denominator d0 q, and the near-singular m = 0 term must keep its
    accuracy through it (per-vortex gate of the module docstring; mixed 1e-5)."""
    pop = wl.periodic(300, 11, box_length=2.0)
    k = kernel_for("periodic", precision, boxLengthX=2.0, boxLengthY=2.0)
    k.set_launch(symmetric=symmetric)
    for sep in seps:
        x, y = pop.x.copy(), pop.y.copy()
        x[:2] = [0.0, 0.8 * sep]
        y[:2] = [0.0, 0.6 * sep]
        vel = k.evaluate(VortexSystem(x, y, pop.circulation))
        ur, vr = oracle.evaluate(spec("periodic", box_length=2.0), x, y, pop.circulation)
        assert per_vortex(vel.x, vel.y, ur, vr) <= bound, sep


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
def test_hamiltonian_gradient_and_conservation(geometry):
    """tests/audit_tests.cpp:51-85: velocity = (1/G)(dH/dy, -dH/dx) by central differences, and
    H conserved over 100 RK4 steps at dt = 1e-4."""
    kw = {"coreRadius": 0.05} if geometry == "infinite" else {}
    if geometry == "periodic":
        kw = {"boxLengthX": 2.0, "boxLengthY": 2.0}
    k = kernel_for(geometry, **kw)
    s = VortexSystem([-.31, .23, .11, -.14], [-.12, .14, -.29, .32], [1, -1, 2, -2])
    vel = k.evaluate(s)
    h = 1e-6
    for i in range(4):
        p, m = s.copy(), s.copy()
        p.x[i] += h
        m.x[i] -= h
        dx = (k.hamiltonian(p) - k.hamiltonian(m)) / (2 * h)
        p, m = s.copy(), s.copy()
        p.y[i] += h
        m.y[i] -= h
        dy = (k.hamiltonian(p) - k.hamiltonian(m)) / (2 * h)
        assert abs(vel.x[i] - dy / s.circulation[i]) <= 2e-8
        assert abs(vel.y[i] + dx / s.circulation[i]) <= 2e-8
    h0 = k.hamiltonian(s)
    k.uploadDeviceState(s)
    for _ in range(100):
        k.deviceRk4Step(1e-4)
    e = s.copy()
    k.downloadDeviceState(e)
    assert abs(k.hamiltonian(e) - h0) <= 1e-10


def test_invalid_inputs():
    """tests/audit_tests.cpp:99-118 and the reference's validation messages."""
    k = kernel_for("infinite")
    with pytest.raises(InvalidArgument, match="vortex data must be finite"):
        k.evaluate(VortexSystem([np.nan], [0.0], [1.0]))
    with pytest.raises(InvalidArgument, match="timestep must be finite and positive"):
        k.uploadDeviceState(VortexSystem([0.0], [0.0], [1.0]))
        k.deviceRk4Step(0.0)
    with pytest.raises(OutOfRange, match="invalid target-vortex range"):
        k.evaluateRange([0.0, 1.0], [0.0, 0.0], [1.0, 1.0], 1, 3)
    with pytest.raises(BackendRuntimeError,
                       match="coincident vortices in singular point-vortex kernel"):
        k.evaluate(VortexSystem([0.5, 0.5, 0.1], [0.2, 0.2, 0.3], [1.0, -1.0, 1.0]))
    p = kernel_for("periodic")
    with pytest.raises(InvalidArgument, match="periodic box requires zero total circulation"):
        p.evaluate(VortexSystem([0.0, 0.5], [0.0, 0.1], [1.0, 1.0]))
    d = kernel_for("disk")
    with pytest.raises(InvalidArgument, match="vortex lies on or outside the disk"):
        d.evaluate(VortexSystem([1.0, 0.0], [0.0, 0.1], [1.0, 1.0]))
    with pytest.raises(BackendRuntimeError, match="coincident vortices in disk"):
        d.evaluate(VortexSystem([0.1, 0.1], [0.2, 0.2], [1.0, 1.0]))
    fresh = kernel_for("infinite")
    with pytest.raises(LogicError):
        fresh.deviceRk4Step(1e-3)
    with pytest.raises(InvalidArgument):
        kernel_for("infinite", coreRadius=-1.0)
    with pytest.raises(InvalidArgument, match="square periodic box"):
        kernel_for("periodic", boxLengthX=2.0, boxLengthY=3.0)


# ------------------------------------------------------------------ seeded parity


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
@pytest.mark.parametrize("n", [2, 4, 258, 1000, 4096])
def test_fp64_velocity_parity(oracle, geometry, n):
    pop = population(geometry, n, seed=11 + n)
    k = kernel_for(geometry, boxLengthX=2.0, boxLengthY=2.0)
    vel = k.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
    ur, vr = oracle.evaluate(spec_for(geometry, boxLengthX=2.0), pop.x, pop.y, pop.circulation)
    gate = fp64_gate(geometry, pop.x, pop.y, pop.circulation, ur, vr)
    assert normwise(vel.x, vel.y, ur, vr) <= gate
    assert per_vortex(vel.x, vel.y, ur, vr) <= 100 * gate
    if geometry == "disk":
        # Against the extended-precision sum itself, where the oracle's own FP64 loses digits.
        ue, ve = disk_truth(pop.x, pop.y, pop.circulation)
        assert normwise(vel.x, vel.y, ue, ve) <= 1e-12


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
@pytest.mark.parametrize("n", [258, 4096])
def test_mixed_velocity_parity(oracle, geometry, n):
    pop = population(geometry, n, seed=5 + n)
    k = kernel_for(geometry, "mixed", boxLengthX=2.0, boxLengthY=2.0)
    vel = k.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
    ur, vr = oracle.evaluate(spec_for(geometry, boxLengthX=2.0), pop.x, pop.y, pop.circulation)
    assert normwise(vel.x, vel.y, ur, vr) <= 1e-6


@pytest.mark.parametrize("precision,bound", [("fp64", 1e-12), ("mixed", 1e-6)])
@pytest.mark.parametrize("layers", [0, 1, 2, 3, 6, 8, 20])
def test_periodic_layers_parity(oracle, layers, precision, bound):
    """Rows m = +-1 are evaluated in closed form and rows |m| >= 2 as a series: L = 0, 1, 2
    switch each part on in turn."""
    pop = wl.periodic(300, 77, box_length=TWO_PI)
    k = kernel_for("periodic", precision, boxLengthX=TWO_PI, boxLengthY=TWO_PI,
                   periodicImageLayers=layers)
    vel = k.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
    ur, vr = oracle.evaluate(spec("periodic", box_length=TWO_PI, image_layers=layers), pop.x,
                             pop.y, pop.circulation)
    assert normwise(vel.x, vel.y, ur, vr) <= bound


def test_periodic_positions_outside_box(oracle):
    """The reference never wraps positions; remainder() reduces every difference."""
    pop = wl.periodic(200, 9, box_length=2.0)
    x = pop.x + 2.0 * np.random.default_rng(1).integers(-3, 4, 200)
    k = kernel_for("periodic", boxLengthX=2.0, boxLengthY=2.0)
    vel = k.evaluate(VortexSystem(x, pop.y, pop.circulation))
    ur, vr = oracle.evaluate(spec("periodic", box_length=2.0), x, pop.y, pop.circulation)
    assert normwise(vel.x, vel.y, ur, vr) <= 1e-12


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
def test_evaluate_range_subsets(oracle, geometry):
    pop = population(geometry, 700, seed=3)
    k = kernel_for(geometry, boxLengthX=2.0, boxLengthY=2.0)
    s = spec_for(geometry, boxLengthX=2.0)
    for b, e in [(0, 1), (13, 300), (650, 700), (0, 700), (5, 5)]:
        vel = k.evaluateRange(pop.x, pop.y, pop.circulation, b, e)
        ur, vr = oracle.evaluate_range(s, pop.x, pop.y, pop.circulation, b, e)
        if e > b:
            assert normwise(vel.x[b:e], vel.y[b:e], ur[b:e], vr[b:e]) <= 1e-12
        assert np.all(vel.x[:b] == 0.0) and np.all(vel.x[e:] == 0.0)


@pytest.mark.parametrize("tpt,splits", [(1, 1), (1, 3), (2, 0), (4, 0), (4, 7), (2, 64)])
def test_launch_shapes_agree(oracle, tpt, splits):
    pop = wl.disk(3000, 21)
    k = kernel_for("disk")
    k.set_launch(tpt, splits)
    vel = k.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
    ur, vr = oracle.evaluate(spec("disk"), pop.x, pop.y, pop.circulation)
    assert normwise(vel.x, vel.y, ur, vr) <= 1e-12


def test_deterministic_repeats():
    pop = wl.periodic(2000, 4, box_length=2.0)
    k = kernel_for("periodic", boxLengthX=2.0, boxLengthY=2.0)
    s = VortexSystem(pop.x, pop.y, pop.circulation)
    a, b = k.evaluate(s), k.evaluate(s)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


# ------------------------------------------------------------------ stepping


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
def test_rk4_steps_parity(oracle, geometry):
    pop = population(geometry, 500, seed=8)
    k = kernel_for(geometry, boxLengthX=2.0, boxLengthY=2.0)
    k.uploadDeviceState(VortexSystem(pop.x, pop.y, pop.circulation))
    dt = 1e-4 if geometry == "disk" else 1e-3
    for _ in range(10):
        k.deviceRk4Step(dt)
    out = VortexSystem(pop.x.copy(), pop.y.copy(), pop.circulation)
    k.downloadDeviceState(out)
    it = oracle.integrator(500)
    x, y = pop.x.copy(), pop.y.copy()
    s = spec_for(geometry, boxLengthX=2.0)
    for _ in range(10):
        it.rk4_step(s, x, y, pop.circulation, dt)
    scale = np.max(np.abs(np.concatenate([x, y])))
    assert np.max(np.abs(out.x - x)) <= 1e-12 * scale
    assert np.max(np.abs(out.y - y)) <= 1e-12 * scale


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reference_periodic_state(precision):
    """A state the reference keeps among its plotting examples: 200 vortices of +-0.005 in the
    2 pi box, 8 image layers, against the reference's own outputs (tests/golden): velocities
    (1e-12 FP64, 1e-6 mixed); in FP64 also 10 RK4 steps at dt = 1e-2 (1e-12 of the coordinate
    scale) and a DOPRI5 run to t = 0.5 with its accepted-step records (the same steps,
    accepted dt to 1e-7 relative, and the positions to 1e-10 of the coordinate scale)."""
    from pathlib import Path

    d = np.load(Path(__file__).resolve().parent / "golden" / "reference_periodic_state.npz")
    x0, y0, g = d["x"], d["y"], d["g"]
    kw = dict(boxLengthX=TWO_PI, boxLengthY=TWO_PI, periodicImageLayers=8)
    k = kernel_for("periodic", precision, **kw)
    vel = k.evaluate(VortexSystem(x0, y0, g))
    assert normwise(vel.x, vel.y, d["u"], d["v"]) <= (1e-12 if precision == "fp64" else 1e-6)
    if precision == "mixed":
        return
    scale = np.max(np.abs(np.concatenate([x0, y0])))
    k.uploadDeviceState(VortexSystem(x0, y0, g))
    for _ in range(10):
        k.deviceRk4Step(float(d["rk4_dt"][0]))
    out = k.downloadDeviceSystem()
    assert np.max(np.abs(out.x - d["rk4_x"])) <= 1e-12 * scale
    assert np.max(np.abs(out.y - d["rk4_y"])) <= 1e-12 * scale
    k = kernel_for("periodic", **kw)
    k.uploadDeviceState(VortexSystem(x0, y0, g))
    rec = d["dopri_record"]
    time, dt = 0.0, 1e-2
    for row in rec:
        r = k.deviceDopri5Step(min(dt, 0.5 - time), 1e-12, 1e-9, 1e-12, 0.1)
        assert r.rejectedSteps == int(row[3])
        assert abs(r.acceptedTimeStep - row[0]) <= 1e-7 * row[0]
        time += r.acceptedTimeStep
        dt = r.suggestedTimeStep
    assert abs(time - 0.5) <= 1e-12
    out = k.downloadDeviceSystem()
    assert np.max(np.abs(out.x - d["dopri_x"])) <= 1e-10 * scale
    assert np.max(np.abs(out.y - d["dopri_y"])) <= 1e-10 * scale


def test_rk4_advance_matches_single_steps():
    pop = wl.infinite(1024, 0)
    a = kernel_for("infinite")
    b = kernel_for("infinite")
    s = VortexSystem(pop.x, pop.y, pop.circulation)
    a.uploadDeviceState(s)
    b.uploadDeviceState(s)
    for _ in range(7):
        a.deviceRk4Step(1e-3)
    b.deviceRk4Advance(1e-3, 7)
    ra, rb = s.copy(), s.copy()
    a.downloadDeviceState(ra)
    b.downloadDeviceState(rb)
    assert np.array_equal(ra.x, rb.x) and np.array_equal(ra.y, rb.y)


def test_rk4_pair_rotation():
    """tests/tests.cpp:42-51: co-rotating pair to t=1 at dt=1e-3 (2e-12)."""
    k = kernel_for("infinite")
    s = VortexSystem([-1.0, 1.0], [0.0, 0.0], [1.0, 1.0])
    k.uploadDeviceState(s)
    k.deviceRk4Advance(1e-3, 1000)
    k.downloadDeviceState(s)
    ang = 1.0 / (4.0 * np.pi)
    assert abs(s.x[1] - np.cos(ang)) <= 2e-12 and abs(s.y[1] - np.sin(ang)) <= 2e-12


def test_dopri_pair_rotation():
    """tests/tests.cpp:52-70: DOPRI5 to t=1 (2e-10)."""
    k = kernel_for("infinite")
    s = VortexSystem([-1.0, 1.0], [0.0, 0.0], [1.0, 1.0])
    k.uploadDeviceState(s)
    t, dt = 0.0, 0.01
    while t < 1.0:
        r = k.deviceDopri5Step(min(dt, 1.0 - t), 1e-12, 1e-10, 1e-12, 0.2)
        t += r.acceptedTimeStep
        dt = r.suggestedTimeStep
    k.downloadDeviceState(s)
    ang = 1.0 / (4.0 * np.pi)
    assert abs(s.x[1] - np.cos(ang)) <= 2e-10 and abs(s.y[1] - np.sin(ang)) <= 2e-10


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
def test_dopri5_parity(oracle, geometry):
    pop = population(geometry, 300, seed=12)
    k = kernel_for(geometry, boxLengthX=2.0, boxLengthY=2.0)
    k.uploadDeviceState(VortexSystem(pop.x, pop.y, pop.circulation))
    it = oracle.integrator(300)
    s = spec_for(geometry, boxLengthX=2.0)
    x, y = pop.x.copy(), pop.y.copy()
    dt_g = dt_o = 1e-3
    for _ in range(6):
        rg = k.deviceDopri5Step(dt_g, 1e-12, 1e-9, 1e-12, 0.1)
        ro = it.dopri5_step(s, x, y, pop.circulation, dt_o, 1e-12, 1e-9, 1e-12, 0.1)
        # The error norm is a difference of stage velocities, so the ~1e-14 relative velocity
        # differences surface there amplified; step sizes follow it through err^-0.2.
        assert rg.rejectedSteps == ro[3]
        assert abs(rg.acceptedTimeStep - ro[0]) <= 1e-7 * ro[0]
        assert abs(rg.normalizedError - ro[2]) <= 1e-5 * max(ro[2], 1e-300)
        dt_g, dt_o = rg.suggestedTimeStep, ro[1]
    out = VortexSystem(x.copy(), y.copy(), pop.circulation)
    k.downloadDeviceState(out)
    scale = np.max(np.abs(np.concatenate([x, y])))
    assert np.max(np.abs(out.x - x)) <= 1e-10 * scale
    assert np.max(np.abs(out.y - y)) <= 1e-10 * scale


def test_dopri_fsal_counts_and_invalidation(oracle):
    """tests/tests.cpp:297-314: the co-rotating pair takes 7 evaluations in its first DOPRI5
    step, 13 after the second (the FSAL stage is reused) and 20 after an invalidation (the
    first stage is recomputed), counted by the handle's launch counter.  On a 64-vortex
    population the same bookkeeping holds with rejections (6 more per rejected attempt), and
    the errors match the oracle's integrator driven the same way."""
    k = kernel_for("infinite")
    k.uploadDeviceState(VortexSystem([-1.0, 1.0], [0.0, 0.0], [1.0, 1.0]))
    k.profileBegin()
    counts = []
    for step in range(3):
        if step == 2:
            k.invalidateDeviceDerivative()
        r = k.deviceDopri5Step(1e-3, 1e-10, 1e-8, 1e-12, 0.1)
        assert r.rejectedSteps == 0
        counts.append(k.profileEnd()[1])
        k.profileBegin()
    k.profileEnd()
    assert np.cumsum(counts).tolist() == [7, 13, 20]

    pop = wl.infinite(64, 3)
    k = kernel_for("infinite")
    k.uploadDeviceState(VortexSystem(pop.x, pop.y, pop.circulation))
    it = oracle.integrator(64)
    x, y = pop.x.copy(), pop.y.copy()
    s = spec("infinite")
    for step in range(3):
        if step == 2:
            k.invalidateDeviceDerivative()
            it.invalidate()
        k.profileBegin()
        rg = k.deviceDopri5Step(1e-3, 1e-10, 1e-8, 1e-12, 0.1)
        launches = k.profileEnd()[1]
        ro = it.dopri5_step(s, x, y, pop.circulation, 1e-3, 1e-10, 1e-8, 1e-12, 0.1)
        assert rg.rejectedSteps == ro[3]
        assert launches == (step != 1) + 6 * (1 + rg.rejectedSteps)
        assert abs(rg.normalizedError - ro[2]) <= 1e-6 * ro[2]


def test_dopri_nonfinite_and_empty():
    k = kernel_for("infinite")
    k.uploadDeviceState(VortexSystem([], [], []))
    r = k.deviceDopri5Step(0.01, 1e-10, 1e-8, 1e-12, 0.1)
    assert r.acceptedTimeStep == 0.01 and r.suggestedTimeStep == 0.05
    k.uploadDeviceState(VortexSystem([0.0, 0.0], [0.0, 0.0], [1.0, 1.0]))
    with pytest.raises(BackendRuntimeError):
        k.deviceDopri5Step(0.01, 1e-10, 1e-8, 1e-12, 0.1)


# ------------------------------------------------------------------ invariants, dipoles


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
@pytest.mark.parametrize("n", [3, 500, 2048])
def test_invariants_parity(oracle, geometry, n):
    if geometry == "periodic" and n % 2:
        n += 1
    pop = population(geometry, n, seed=40 + n)
    kw = {"coreRadius": 0.01} if geometry == "infinite" else {}
    k = kernel_for(geometry, boxLengthX=2.0, boxLengthY=2.0, **kw)
    inv = k.invariantsOf(VortexSystem(pop.x, pop.y, pop.circulation))
    ref = oracle.invariants(spec_for(geometry, boxLengthX=2.0, **kw), pop.x, pop.y,
                            pop.circulation)
    got = np.array([inv.circulation, inv.linearImpulseX, inv.linearImpulseY, inv.angularImpulse,
                    inv.hamiltonian])
    scale = np.sum(np.abs(pop.circulation)) * max(1.0, np.max(np.abs(pop.x)) ** 2)
    assert np.all(np.abs(got[:4] - ref[:4]) <= 1e-13 * scale)
    # H against its extended-precision value: every FP64 term to a few ulps and a compensated
    # sum leave at most a small multiple of 2^-53 S (S: the terms' magnitudes, see
    # hamiltonian_longdouble); against the oracle, that plus the oracle's own distance.
    h_true, s_abs = hamiltonian_longdouble(geometry, pop.x, pop.y, pop.circulation,
                                           core_radius=kw.get("coreRadius", 0.0))
    bar = 16 * 2.0 ** -53 * s_abs
    assert abs(got[4] - h_true) <= bar, (got[4], h_true, bar)
    assert abs(got[4] - ref[4]) <= abs(ref[4] - h_true) + bar
    k.uploadDeviceState(VortexSystem(pop.x, pop.y, pop.circulation))
    dev = k.deviceInvariants()
    assert dev.hamiltonian == inv.hamiltonian


@pytest.mark.parametrize("layers", [0, 1, 2, 3, 8, 20])
def test_periodic_hamiltonian_against_the_reference_build(reference, layers):
    """The periodic Hamiltonian is this backend's own form (rows m != 0 as a Fourier series);
    the reference sums the 2L+1 rows directly (src/compute.cpp:174-202).  At every L: within
    16 * 2^-53 S of the extended-precision row sum, and against the reference build itself
    within that plus the reference's own distance from it.  Two box lengths."""
    for box, seed in ((2.0, 1), (TWO_PI, 2)):
        pop = wl.periodic(400, 60 + layers + seed, box_length=box)
        k = kernel_for("periodic", boxLengthX=box, boxLengthY=box, periodicImageLayers=layers)
        h = k.hamiltonian(VortexSystem(pop.x, pop.y, pop.circulation))
        sp = spec("periodic", box_length=box, image_layers=layers)
        h_ref = reference.invariants(sp, pop.x, pop.y, pop.circulation)[4]
        h_true, s_abs = hamiltonian_longdouble("periodic", pop.x, pop.y, pop.circulation,
                                               box_length=box, layers=layers)
        bar = 16 * 2.0 ** -53 * s_abs
        assert abs(h - h_true) <= bar, (box, h, h_true, bar)
        assert abs(h - h_ref) <= abs(h_ref - h_true) + bar, (box, h, h_ref)


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
def test_dipole_candidates_parity(oracle, geometry):
    pop = population(geometry, 3000, seed=90)
    if geometry == "disk":
        pop.circulation[:] = np.where(np.arange(3000) % 2 == 0, 1.0, -1.0)
    dist = 0.02
    k = kernel_for(geometry, boxLengthX=2.0, boxLengthY=2.0)
    k.uploadDeviceState(VortexSystem(pop.x, pop.y, pop.circulation))
    f, s, d2 = k.dipoleCandidates(dist)
    fo, so, d2o = oracle.dipole_candidates(geometry, 2.0, 1.0, dist, pop.x, pop.y,
                                           pop.circulation)
    assert len(fo) > 0
    got = sorted(zip(d2.tolist(), f.tolist(), s.tolist()))
    want = sorted(zip(d2o.tolist(), fo.tolist(), so.tolist()))
    assert got == want


@pytest.mark.parametrize("precision,gate", [("fp64", None), ("mixed", 1e-6)])
def test_disk_near_centre_source_in_a_population(oracle, precision, gate):
    """A source within 2^-100 R of the centre puts its tile on the reference's a, b form for
    the image (its image point would overflow); the other tiles use the image-point form."""
    pop = wl.disk(1200, 41)
    x, y = pop.x.copy(), pop.y.copy()
    x[700], y[700] = 3e-160, -1e-161      # one near-centre source (two would nearly coincide)
    k = kernel_for("disk", precision)
    vel = k.evaluate(VortexSystem(x, y, pop.circulation))
    ur, vr = oracle.evaluate(spec("disk"), x, y, pop.circulation)
    if gate is None:
        gate = fp64_gate("disk", x, y, pop.circulation, ur, vr)
    assert normwise(vel.x, vel.y, ur, vr) <= gate


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
def test_fp64_parity_at_n65536_on_target_subsets(oracle, geometry):
    """The benchmark's size: the full N=65536 evaluation on the device against the oracle's
    sums for 768 targets (the first 256, 256 from the middle and the last 256, which for the
    disk are the wall shell at r in [0.995, 0.9999]) over all sources.  For the disk the bar
    against the oracle is max(1e-12, 2x the oracle's own distance from an extended-precision
    sum on the same targets), and against that sum itself 1e-12."""
    n = 65536
    if geometry == "infinite":
        pop = wl.infinite(n, 0)
    elif geometry == "periodic":
        pop = wl.periodic(n, 1)
    else:
        pop = wl.disk(n, 2)
    kw = {"boxLengthX": TWO_PI, "boxLengthY": TWO_PI}
    k = kernel_for(geometry, **kw)
    vel = k.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
    s = spec_for(geometry, boxLengthX=TWO_PI)
    for b, e in [(0, 256), (n // 2, n // 2 + 256), (n - 256, n)]:
        ur, vr = oracle.evaluate_range(s, pop.x, pop.y, pop.circulation, b, e)
        ur, vr = ur[b:e], vr[b:e]
        gate = 1e-12
        if geometry == "disk":
            from tests.conftest import disk_longdouble_targets
            ue, ve = disk_longdouble_targets(pop.x, pop.y, pop.circulation, np.arange(b, e))
            gate = max(gate, 2.0 * normwise(ur, vr, ue, ve))
            # The wall shell (the last 256 targets) against the extended-precision sum.
            assert normwise(vel.x[b:e], vel.y[b:e], ue, ve) <= 1e-12, (b, e)
        assert normwise(vel.x[b:e], vel.y[b:e], ur, vr) <= gate, (b, e)


@pytest.mark.parametrize("geometry", ["infinite", "periodic", "disk"])
def test_mixed_parity_at_n65536_on_target_subsets(oracle, geometry):
    """Mixed precision at the benchmark's size against the FP64 oracle, on the same target
    subsets (normwise 1e-6)."""
    n = 65536
    pop = {"infinite": lambda: wl.infinite(n, 0), "periodic": lambda: wl.periodic(n, 1),
           "disk": lambda: wl.disk(n, 2)}[geometry]()
    k = kernel_for(geometry, "mixed", boxLengthX=TWO_PI, boxLengthY=TWO_PI)
    vel = k.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
    s = spec_for(geometry, boxLengthX=TWO_PI)
    for b, e in [(0, 256), (n - 256, n)]:
        ur, vr = oracle.evaluate_range(s, pop.x, pop.y, pop.circulation, b, e)
        assert normwise(vel.x[b:e], vel.y[b:e], ur[b:e], vr[b:e]) <= 1e-6, (b, e)


def test_mixed_disk_run_against_the_oracle_at_first_and_last_step(oracle):
    """The N=65536 disk RK4 run of the benchmark (dt=1e-4, 200 steps) in mixed precision: the
    mixed kernel's velocities on the run's first and last states against the oracle's FP64
    sums for 512 targets (the first 256 and the wall shell, the last 256), normwise 1e-6."""
    n = 65536
    pop = wl.disk(n, 2)
    k = kernel_for("disk", "mixed")
    s = VortexSystem(pop.x, pop.y, pop.circulation)
    k.uploadDeviceState(s)
    states = [s.copy()]
    k.deviceRk4Advance(1e-4, 200)
    last = s.copy()
    k.downloadDeviceState(last)
    states.append(last)
    for st in states:
        assert np.all(np.hypot(st.x, st.y) < 1.0)
        vel = k.evaluate(VortexSystem(st.x, st.y, pop.circulation))
        for b, e in [(0, 256), (n - 256, n)]:
            ur, vr = oracle.evaluate_range(spec("disk"), st.x, st.y, pop.circulation, b, e)
            assert normwise(vel.x[b:e], vel.y[b:e], ur[b:e], vr[b:e]) <= 1e-6, (b, e)


def test_mixed_disk_self_image_stays_out_of_the_fp32_sums(oracle):
    """128 vortices uniform in the unit disk, one of them 5.5e-4 from the wall: its velocity
    (about 146, against an rms of 15) is nearly all its own image.  Inside an FP32 tile sum
    that term made every later FP32 addition round at its scale, 1.0e-6 normwise against FP64
    on this population.  The self-image term is added in FP64 where each target's velocity is
    finalized, so every path that evaluates the mixed disk stays well inside the bar: the
    split and symmetric single-system paths, both ensemble decompositions, and one RK4 step
    of the persistent kernel (displacement against the oracle's step)."""
    from pointvortex_amd.ensemble import Ensemble, InstanceConfig

    pop = wl.disk(128, 17 * 1_000_003 + 1083, wall_fraction=0.0, gaussian=False)
    r = np.hypot(pop.x, pop.y)
    assert 5e-4 < 1.0 - r.max() < 6e-4
    ur, vr = oracle.evaluate(spec("disk"), pop.x, pop.y, pop.circulation)
    bar = 5e-7
    for shape in ((1, 1, 128, 0), (1, 0, 128, 1), (2, 0, 128, 1)):
        k = kernel_for("disk", "mixed")
        k.set_launch(*shape)
        vel = k.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
        assert normwise(vel.x, vel.y, ur, vr) <= bar, shape
    for symmetric in (0, 1):
        e = Ensemble(2 * pop.n, 2, "mixed")
        e.set_symmetric(symmetric)
        cfg = InstanceConfig("disk", disk_radius=1.0)
        e.load([cfg, cfg], [pop.x, pop.x], [pop.y, pop.y], [pop.circulation] * 2)
        us, vs = e.evaluate()
        e.close()
        for q in range(2):
            assert normwise(us[q], vs[q], ur, vr) <= bar, (symmetric, q)
    dt = 1e-6
    x, y = pop.x.copy(), pop.y.copy()
    oracle.integrator(pop.n).rk4_step(spec("disk"), x, y, pop.circulation, dt)
    k = kernel_for("disk", "mixed")
    s = VortexSystem(pop.x, pop.y, pop.circulation)
    k.uploadDeviceState(s)
    k.deviceRk4Step(dt)
    k.downloadDeviceState(s)
    assert normwise(s.x - pop.x, s.y - pop.y, x - pop.x, y - pop.y) <= bar


def _planted(geometry, seed):
    """300 vortices with four opposite-sign pairs moved to within 3e-4 of each other (and, in
    the disk, three vortices within 2e-4 of the wall)."""
    if geometry == "periodic":
        pop = wl.periodic(300, seed, box_length=1.0)
    elif geometry == "disk":
        pop = wl.disk(300, seed, wall_fraction=0.0, gaussian=False)
    else:
        pop = wl.infinite(300, seed)
    x, y, g = pop.x.copy(), pop.y.copy(), pop.circulation.copy()
    rng = np.random.default_rng(seed)
    used = set()
    for i in rng.permutation(300)[:40]:
        if len(used) >= 8 or i in used:
            continue
        j = next(k for k in range(300) if k not in used and k != i and g[k] == -g[i])
        x[j], y[j] = x[i] + 3e-4 * (1 + 0.1 * len(used)), y[i]
        used |= {i, j}
    if geometry == "disk":
        free = [k for k in range(300) if k not in used][:3]
        for k, a in zip(free, (0.3, 2.0, 4.0)):
            x[k], y[k] = (1 - 1.5e-4) * np.cos(a), (1 - 1.5e-4) * np.sin(a)
    return x, y, g


@pytest.mark.parametrize("geometry,mode", [("periodic", "independent"), ("periodic", "paired"),
                                           ("disk", "paired"), ("disk", "none"),
                                           ("infinite", "none")])
def test_dipole_process_matches_reference(reference, geometry, mode):
    """DipoleManager::process on the device state (device search, host selection, compaction
    and seeded reinjection) against the reference's own, bit for bit."""
    x, y, g = _planted(geometry, 31)
    modes = {"none": 0, "independent": 1, "paired": 2}
    box = 1.0 if geometry == "periodic" else 2.0
    rx, ry, rg, rev = reference.dipole_process(geometry, box, 1.0, 1e-3, modes[mode], 42, x, y, g)
    k = kernel_for(geometry, boxLengthX=box, boxLengthY=box)
    k.uploadDeviceState(VortexSystem(x, y, g))
    k.setDipoleRandom(42)
    ev = k.processDipoles(1e-3, mode)
    assert ev == rev and ev >= 4
    out = k.downloadDeviceSystem()
    assert out.size() == len(rx)
    assert np.array_equal(out.x, rx) and np.array_equal(out.y, ry)
    assert np.array_equal(out.circulation, rg)


def test_dipole_known_answers():
    """The reference's dipole unit tests (tests/tests.cpp:169-296) on the device path, with
    the default generator seed 1234567 (src/params.h) unless the case sets one."""
    def run(geometry, x, y, g, distance, mode, seed=1234567, box=2.0):
        k = kernel_for(geometry, boxLengthX=box, boxLengthY=box, diskRadius=1.0)
        k.uploadDeviceState(VortexSystem(x, y, g))
        k.setDipoleRandom(seed)
        return k.processDipoles(distance, mode), k.downloadDeviceSystem()

    # Closest-first removal: the (0, 0.01) pair goes, the vortex at 0.02 stays.
    ev, s = run("infinite", [0.0, 0.01, 0.02], [0.0, 0.0, 0.0], [1.0, -1.0, -1.0], 0.05, "none")
    assert ev == 1 and s.size() == 1 and s.x[0] == 0.02
    # Periodic minimum image across the box edge, paired reinjection inside the box.
    ev, s = run("periodic", [0.99, -0.99, -0.5, 0.5], [0.0, 0.0, 0.5, -0.5],
                [1.0, -1.0, 1.0, -1.0], 0.05, "paired", seed=42)
    assert ev == 1 and s.size() == 4
    assert np.all(np.abs(s.x) <= 1.0) and np.all(np.abs(s.y) <= 1.0)
    dx = np.remainder(s.x[2] - s.x[3] + 1.0, 2.0) - 1.0
    dy = np.remainder(s.y[2] - s.y[3] + 1.0, 2.0) - 1.0
    assert abs(dx * dx + dy * dy - 1.0) <= 2e-15
    # Disk, independent reinjection keeps the circulations and stays inside the disk.
    ev, s = run("disk", [0.0, 0.01], [0.0, 0.0], [2.0, -0.5], 0.1, "independent")
    assert s.size() == 2 and list(s.circulation) == [2.0, -0.5]
    assert np.all(s.x * s.x + s.y * s.y < 1.0)
    # Disk, paired reinjection: the pair comes back at squared spacing pi / 2.
    ev, s = run("disk", [0.0, 0.01], [0.0, 0.0], [1.0, -1.0], 0.1, "paired")
    d2 = (s.x[0] - s.x[1]) ** 2 + (s.y[0] - s.y[1]) ** 2
    assert abs(d2 - np.pi / 2.0) <= 3e-15
    assert np.all(s.x * s.x + s.y * s.y < 1.0)
    # A vortex 0.02 from the wall forms a dipole with its own image and is removed; with
    # paired reinjection the population keeps its size and circulation, inside the disk.
    ev, s = run("disk", [0.98, 0.0], [0.0, 0.0], [1.0, -1.0], 0.05, "none")
    assert ev == 1 and s.size() == 1 and s.circulation[0] == -1.0
    ev, s = run("disk", [0.98], [0.0], [2.0], 0.05, "paired")
    assert ev == 1 and s.size() == 1 and s.circulation[0] == 2.0
    assert s.x[0] ** 2 + s.y[0] ** 2 < 1.0


def test_dipole_generator_state_round_trip(reference):
    """The generator's text state (as checkpoints store it) continues the sequence: a second
    handle restored from it reinjects exactly as the first handle does next."""
    x, y, g = _planted("periodic", 32)
    k = kernel_for("periodic", boxLengthX=1.0, boxLengthY=1.0)
    k.setDipoleRandom(7)
    k.uploadDeviceState(VortexSystem(x, y, g))
    k.processDipoles(1e-3, "independent")
    state = k.dipoleRandomState()
    x2, y2, g2 = _planted("periodic", 33)
    k.uploadDeviceState(VortexSystem(x2, y2, g2))
    k.processDipoles(1e-3, "independent")
    a = k.downloadDeviceSystem()
    q = kernel_for("periodic", boxLengthX=1.0, boxLengthY=1.0)
    q.setDipoleRandom(7, state)
    q.uploadDeviceState(VortexSystem(x2, y2, g2))
    q.processDipoles(1e-3, "independent")
    b = q.downloadDeviceSystem()
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    assert np.array_equal(a.circulation, b.circulation)
    with pytest.raises(BackendRuntimeError, match="random-generator state"):
        q.setDipoleRandom(7, "not a generator state")


def test_dipole_process_then_evaluate(oracle):
    """After an event the resident state is the new one (the FSAL derivative is gone): the
    next evaluation matches the oracle on the compacted, reinjected population."""
    x, y, g = _planted("disk", 34)
    k = kernel_for("disk")
    k.uploadDeviceState(VortexSystem(x, y, g))
    k.setDipoleRandom(3)
    assert k.processDipoles(1e-3, "paired") >= 4
    s = k.downloadDeviceSystem()
    vel = k.evaluateDeviceState()
    ur, vr = oracle.evaluate(spec("disk"), s.x, s.y, s.circulation)
    assert normwise(vel.x, vel.y, ur, vr) <= fp64_gate("disk", s.x, s.y, s.circulation, ur, vr)
