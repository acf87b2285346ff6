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
    assert k.resolved_launch(pop.n, pop.n) == (1, 2, 128)
    k.set_launch(4, 0, 256)                                          # explicit override wins
    assert k.resolved_launch(pop.n, pop.n)[0] == 4 and k.resolved_launch(pop.n, pop.n)[2] == 256
    k.set_launch(0, 0, 0)
    assert k.resolved_launch(pop.n, pop.n) == (1, 2, 128)
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
        assert k.resolved_launch(5000, 5000) == (4, 3, 128)
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
    tpt, splits, tile = res["best"]
    pop = autotune.population(geometry, 2048)
    k = HipKernel(SimParams(boundaryCondition=geometry, precision=precision,
                            boxLengthX=2 * np.pi, boxLengthY=2 * np.pi))
    k.set_launch(tpt, splits, tile)
    vel = k.evaluate(VortexSystem(pop.x, pop.y, pop.circulation))
    ur, vr = oracle.evaluate(spec(geometry, box_length=2 * np.pi), pop.x, pop.y, pop.circulation)
    gate = (fp64_gate(geometry, pop.x, pop.y, pop.circulation, ur, vr) if precision == "fp64"
            else 1e-6)
    assert normwise(vel.x, vel.y, ur, vr) <= gate
