"""Initial-condition files in the reference's format (src/read.cpp:212-280,
initial_conditions/initial_condition.cpp:211-237): exact round trip of the seeded workloads,
the metadata the reference reads, the generator's checks, and the reference's own CPU driver
reading a written file."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from pointvortex_amd import ic
from pointvortex_amd import workloads as wl

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.parametrize("pop,seed", [(wl.infinite(300, 4), 4), (wl.periodic(300, 5), 5),
                                      (wl.disk(300, 6), 6)])
def test_round_trip_is_exact(tmp_path, pop, seed):
    f = ic.write_initial_condition(tmp_path / "a.ic", pop, seed, "test workload")
    back = ic.read_initial_condition(f)
    assert back.geometry == pop.geometry
    assert back.box_length == pop.box_length and back.radius == pop.radius
    for a, b in ((back.x, pop.x), (back.y, pop.y), (back.circulation, pop.circulation)):
        assert np.array_equal(a, b)
    head = f.read_text().splitlines()[:6]
    assert head[0] == "# PointVortex initial condition"
    assert head[1] == f"# geometry={pop.geometry} pattern=random seed={seed}"
    assert head[-1] == "# x y circulation"


def test_minimum_separation_matches_brute_force():
    for pop in (wl.infinite(400, 1), wl.periodic(400, 2, box_length=1.0), wl.disk(400, 3)):
        brute = min(ic._distance(pop.x, pop.y, i, j, pop.geometry, pop.box_length)
                    for i in range(pop.n) for j in range(i))
        assert ic.minimum_separation(pop) == brute


def test_generator_checks_and_no_overwrite(tmp_path):
    pop = wl.periodic(8, 1, box_length=1.0)
    f = ic.write_initial_condition(tmp_path / "p.ic", pop, 1)
    with pytest.raises(FileExistsError):
        ic.write_initial_condition(f, pop, 1)
    ic.write_initial_condition(f, pop, 1, overwrite=True)
    bad = wl.Population("periodic", pop.x + 1.0, pop.y, pop.circulation, box_length=1.0)
    with pytest.raises(ValueError, match="outside the fundamental box"):
        ic.write_initial_condition(tmp_path / "q.ic", bad, 1)
    disk = wl.disk(8, 2)
    disk.x[0], disk.y[0] = 1.0, 0.0
    with pytest.raises(ValueError, match="outside the boundary"):
        ic.write_initial_condition(tmp_path / "d.ic", disk, 2)
    zero = wl.infinite(4, 3)
    zero.circulation[1] = 0.0
    with pytest.raises(ValueError, match="invalid vortex"):
        ic.write_initial_condition(tmp_path / "z.ic", zero, 3)


def test_reader_follows_the_reference_rules(tmp_path):
    f = tmp_path / "r.ic"
    f.write_text("# geometry=disk disk_radius=2.5\n0.5, 0.25, 1  # trailing comment\n\n"
                 "-0.5 0 -1\n")
    p = ic.read_initial_condition(f)
    assert p.geometry == "disk" and p.radius == 2.5 and p.n == 2
    assert list(p.x) == [0.5, -0.5] and list(p.circulation) == [1.0, -1.0]
    f.write_text("1 2 3 4\n")
    with pytest.raises(ValueError, match="line 1"):
        ic.read_initial_condition(f)
    f.write_text("# only comments\n")
    with pytest.raises(ValueError, match="empty"):
        ic.read_initial_condition(f)


def test_c1_file_matches_the_workload(tmp_path):
    f = ic.write_config("c1", tmp_path)
    p = ic.read_initial_condition(f)
    ref = wl.config1()
    assert np.array_equal(p.x, ref.x) and np.array_equal(p.circulation, ref.circulation)


CPU_DRIVER = ROOT / "oracle" / "_ref" / "point_vortex_cpu"


@pytest.mark.skipif(not CPU_DRIVER.exists(), reason="reference CPU driver not built")
def test_reference_driver_reads_a_written_file(tmp_path):
    """The reference's CPU driver runs from a written file, and checks its metadata against
    the params (src/main.cpp:75-90): a box length that disagrees is refused."""
    pop = wl.periodic(64, 9, box_length=1.0)
    f = ic.write_initial_condition(tmp_path / "in.ic", pop, 9, "driver check")

    def run(box, name):
        d = tmp_path / name
        params = tmp_path / f"{name}.params"
        params.write_text(f"initialConditionFile {f}\nboundaryCondition periodic\n"
                          f"boxLengthX {box}\nboxLengthY {box}\nendTime 0.002\n"
                          "integrator rk4\ntimeStep 0.001\noutputTime 0.001\n"
                          f"diagnosticsTime 0.001\nnumThreads 1\nrunDirectory {d}\n")
        return d, subprocess.run([str(CPU_DRIVER), str(params)], text=True,
                                 capture_output=True, timeout=60)

    d, r = run(1, "ok")
    assert r.returncode == 0, r.stdout + r.stderr
    assert len((d / "trajectory.csv").read_text().splitlines()) > 64
    _, r = run(2, "mismatch")
    assert r.returncode != 0 and "box" in (r.stdout + r.stderr).lower()
