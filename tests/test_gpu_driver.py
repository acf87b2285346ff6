"""Drop-in test: the reference's driver linked to this backend (oracle/_ref/point_vortex_hip)
against the reference's CPU executable (oracle/_ref/point_vortex_cpu), both built from the
reference sources.  Same case list and tolerances as the reference's backend-consistency
harness (tests/backend_consistency.py:29-69 in the reference): every field of trajectory.csv
and diagnostics.csv must agree within rel 2e-10, abs 2e-11."""
import csv
import math
import subprocess

import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu

ORDINARY = [(-.31, -.12, 1), (.23, .14, -1), (.11, -.29, 2), (-.14, .32, -2)]


def _large():
    rng = np.random.default_rng(1234)
    xy = rng.uniform(-.4, .4, (258, 2))
    return [(float(a), float(b), 1 if i % 2 == 0 else -1) for i, (a, b) in enumerate(xy)]


CASES = (
    [(g + "_run", g, ORDINARY, .003, "rk4") for g in ("infinite", "periodic", "disk")]
    + [(g + "_openmp", g, _large(), 0, "rk4") for g in ("infinite", "periodic", "disk")]
    + [("periodic_close", "periodic", [(0, 0, 1), (1e-10, 0, -1)], 0, "rk4"),
       ("disk_center", "disk", [(.2, 0, 1), (1e-155, 0, 1)], 0, "rk4")]
    + [(g + "_dopri", g, ORDINARY, .003, "dopri5") for g in ("infinite", "periodic", "disk")]
)

# Dipole events through the reference's own host code (src/dipole.cpp): a +-1 pair closer than
# the removal distance is removed and, for the periodic box and the disk, reinjected with the
# driver's seeded generator (src/params.h:64-69).
DIPOLE = ORDINARY + [(.05, .05, 1), (.054, .05, -1)]
DIPOLE_CASES = [
    ("periodic_dipole_independent", "periodic", DIPOLE, .01, "dopri5",
     "dipoleRemoval true\ndipoleRemovalDistance 0.01\ndipoleReinjection independent\n"
     "randomSeed 11\n"),
    ("disk_dipole_paired", "disk", DIPOLE, .01, "rk4",
     "dipoleRemoval true\ndipoleRemovalDistance 0.01\ndipoleReinjection paired\n"
     "randomSeed 5\n"),
    ("infinite_dipole_removed", "infinite", DIPOLE, .01, "rk4",
     "dipoleRemoval true\ndipoleRemovalDistance 0.01\n"),
]


def _run(exe, directory, geometry, rows, end, integrator, extra=""):
    directory.mkdir()
    initial = directory / "initial.dat"
    initial.write_text("".join(f"{x:.17g} {y:.17g} {g:.17g}\n" for x, y, g in rows))
    params = directory / "run.params"
    params.write_text(
        f"initialConditionFile {initial}\nboundaryCondition {geometry}\n"
        f"endTime {end}\nintegrator {integrator}\ntimeStep 0.001\noutputTime 0.002\n"
        "diagnosticsTime 0.001\ncheckpointTime 0.003\nnumThreads 2\n"
        f"runDirectory {directory}\n{extra}")
    r = subprocess.run([str(exe), str(params)], text=True, capture_output=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _compare(a_path, b_path):
    with open(a_path) as f:
        a = list(csv.DictReader(f))
    with open(b_path) as f:
        b = list(csv.DictReader(f))
    assert len(a) == len(b), (a_path, len(a), len(b))
    for ra, rb in zip(a, b):
        assert ra.keys() == rb.keys()
        for key in ra:
            x, y = float(ra[key]), float(rb[key])
            assert math.isfinite(x) and math.isfinite(y)
            assert math.isclose(x, y, rel_tol=2e-10, abs_tol=2e-11), (a_path.name, key, x, y)


@pytest.fixture(scope="module")
def executables():
    cpu, hip = pyoracle.REF_CPU_EXE, pyoracle.REF_HIP_EXE
    if not cpu.exists() or not hip.exists():
        pytest.skip("reference driver builds (oracle/_ref) not present")
    return cpu, hip


@pytest.mark.parametrize("name,geometry,rows,end,integrator", CASES, ids=[c[0] for c in CASES])
def test_driver_matches_cpu_reference(executables, tmp_path, name, geometry, rows, end,
                                      integrator):
    cpu, hip = executables
    _run(cpu, tmp_path / "cpu", geometry, rows, end, integrator)
    out = _run(hip, tmp_path / "hip", geometry, rows, end, integrator)
    assert "backend=HIP" in out
    for fname in ("trajectory.csv", "diagnostics.csv"):
        _compare(tmp_path / "cpu" / fname, tmp_path / "hip" / fname)
    details = (tmp_path / "hip" / "resolved_parameters.txt").read_text()
    assert "hip_arch" in details


def test_driver_reports_reference_errors(executables, tmp_path):
    """A coincident pair fails in both executables with the reference's message."""
    cpu, hip = executables
    msgs = []
    for exe, sub in ((cpu, "cpu"), (hip, "hip")):
        d = tmp_path / sub
        d.mkdir()
        initial = d / "initial.dat"
        initial.write_text("0 0 1\n0 0 -1\n")
        params = d / "run.params"
        params.write_text(f"initialConditionFile {initial}\nboundaryCondition periodic\n"
                          f"endTime 0\nrunDirectory {d}\n")
        r = subprocess.run([str(exe), str(params)], text=True, capture_output=True, timeout=60)
        assert r.returncode == 1
        msgs.append(r.stderr.strip())
    assert msgs[0] == msgs[1] == "error: coincident periodic vortices"


@pytest.mark.parametrize("name,geometry,rows,end,integrator,extra", DIPOLE_CASES,
                         ids=[c[0] for c in DIPOLE_CASES])
def test_driver_dipole_events_match_cpu_reference(executables, tmp_path, name, geometry, rows,
                                                  end, integrator, extra):
    cpu, hip = executables
    _run(cpu, tmp_path / "cpu", geometry, rows, end, integrator, extra)
    out = _run(hip, tmp_path / "hip", geometry, rows, end, integrator, extra)
    assert "backend=HIP" in out
    for fname in ("trajectory.csv", "diagnostics.csv"):
        _compare(tmp_path / "cpu" / fname, tmp_path / "hip" / fname)
    # the event happened (diagnostics.csv counts removed and reinjected pairs)
    with open(tmp_path / "hip" / "diagnostics.csv") as f:
        last = list(csv.DictReader(f))[-1]
    assert int(last["removed_pairs"]) >= 1, name
    assert int(last["reinjected_pairs"]) == (0 if geometry == "infinite" else
                                             int(last["removed_pairs"])), name
