"""The infinite-plane configuration run whole: N=1024, +-1 circulations, RK4 dt=1e-3 to t=1
(1000 steps) on the device, against the CPU oracle.

The run is chaotic (close +-1 pairs are under-resolved at this dt), so positions agree only
over the first steps, and then only as far as rounding allows.  The oracle measures that
itself: twins of its own run, with every initial coordinate moved by one ulp, diverge by
~1e-7..1e-6 after 10 steps, and their Hamiltonian and angular-impulse drifts over the run
spread over a range.  The device run must stay within:

* step 10: max position difference <= 10x the widest twin divergence;
* the whole run: circulation unchanged and linear impulse conserved to 1e-11 (RK4 keeps
  linear invariants and the pair sum is antisymmetric), as in the oracle;
* drifts of H and of the angular impulse inside the twins' range widened by its own width.
"""
import numpy as np
import pytest

from oracle.pyoracle import spec
from pointvortex_amd import HipKernel, SimParams, VortexSystem
from pointvortex_amd import workloads as wl

pytestmark = pytest.mark.gpu

DT = 1e-3
STEPS = 1000


def oracle_run(oracle, k, pop, x, y, steps):
    it = oracle.integrator(len(x))
    x, y = x.copy(), y.copy()
    for _ in range(steps):
        it.rk4_step(k, x, y, pop.circulation, DT)
    return x, y


def test_config1_whole_run(oracle):
    pop = wl.config1()
    k = spec("infinite")
    inv0 = oracle.invariants(k, pop.x, pop.y, pop.circulation)

    kern = HipKernel(SimParams(boundaryCondition="infinite"))
    s = VortexSystem(pop.x, pop.y, pop.circulation)
    kern.uploadDeviceState(s)
    kern.deviceRk4Advance(DT, 10)
    at10 = s.copy()
    kern.downloadDeviceState(at10)
    kern.deviceRk4Advance(DT, STEPS - 10)
    end = s.copy()
    kern.downloadDeviceState(end)
    gpu_inv = kern.deviceInvariants()

    x10, y10 = oracle_run(oracle, k, pop, pop.x, pop.y, 10)
    rng = np.random.default_rng(0)
    twin_d10, drifts = [], []
    xe, ye = oracle_run(oracle, k, pop, pop.x, pop.y, STEPS)
    drifts.append(oracle.invariants(k, xe, ye, pop.circulation) - inv0)
    for _ in range(3):
        tx = np.nextafter(pop.x, pop.x + rng.choice([-1.0, 1.0], pop.n))
        ty = np.nextafter(pop.y, pop.y + rng.choice([-1.0, 1.0], pop.n))
        a, b = oracle_run(oracle, k, pop, tx, ty, 10)
        twin_d10.append(np.max(np.abs(np.concatenate([a - x10, b - y10]))))
        xe, ye = oracle_run(oracle, k, pop, tx, ty, STEPS)
        drifts.append(oracle.invariants(k, xe, ye, pop.circulation) - inv0)
    drifts = np.array(drifts)

    d10 = np.max(np.abs(np.concatenate([at10.x - x10, at10.y - y10])))
    assert d10 <= 10 * max(twin_d10), (d10, twin_d10)

    g = np.array([gpu_inv.circulation, gpu_inv.linearImpulseX, gpu_inv.linearImpulseY,
                  gpu_inv.angularImpulse, gpu_inv.hamiltonian]) - inv0
    assert g[0] == 0.0
    assert abs(g[1]) <= 1e-11 and abs(g[2]) <= 1e-11, g
    assert np.all(np.abs(drifts[:, 1:3]) <= 1e-11), drifts
    for col in (3, 4):            # angular impulse, Hamiltonian
        lo, hi = drifts[:, col].min(), drifts[:, col].max()
        width = hi - lo
        assert lo - width <= g[col] <= hi + width, (col, g[col], drifts[:, col])
