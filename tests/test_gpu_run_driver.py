"""The run driver (pointvortex_amd/lib/point_vortex_hip: this backend's own host driver, dipole
events on the device) against the reference's CPU executable.

* The reference's backend-consistency case list and tolerances (rel 2e-10, abs 2e-11) and the
  dipole-event cases, as for the reference driver linked to this backend.
* The output and restart workflows: schedules, a branch restart, a restart with changed
  intervals, a legacy version-4 checkpoint, overwrite protection, a roundoff end time, tiny
  time scales, a dipole-emptied population and its restart, endTime 0, frame numbering and run
  records.  Each runs under both executables in roots of their own; every run directory is
  then compared whole (CSVs, the checkpoint list and contents, the run record's keys), and the
  times and frames are checked against the schedule itself.
* Checkpoints cross: this driver restarts from the reference's checkpoint and the reference
  from this driver's, each matching the other's own restart.
"""
from pathlib import Path

import pytest

from tests.driver_scenarios import (DRIVER, REFERENCE, checkpoint_times, compare_csv,
                                    compare_runs, frames, params, rows, run, times)
from tests.test_gpu_driver import CASES, DIPOLE_CASES, _run

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not DRIVER.exists() or not REFERENCE.exists(),
                                 reason="driver or reference CPU build not present")]


def approx_list(actual, expected):
    return len(actual) == len(expected) and all(
        abs(a - b) <= 1e-12 * max(abs(b), 1e-300) for a, b in zip(actual, expected))


@pytest.mark.parametrize("name,geometry,rows_,end,integrator", CASES, ids=[c[0] for c in CASES])
def test_run_driver_matches_cpu_reference(tmp_path, name, geometry, rows_, end, integrator):
    _run(REFERENCE, tmp_path / "cpu", geometry, rows_, end, integrator)
    out = _run(DRIVER, tmp_path / "hip", geometry, rows_, end, integrator)
    assert "backend=HIP" in out
    for f in ("trajectory.csv", "diagnostics.csv"):
        compare_csv(tmp_path / "cpu" / f, tmp_path / "hip" / f)


@pytest.mark.parametrize("name,geometry,rows_,end,integrator,extra", DIPOLE_CASES,
                         ids=[c[0] for c in DIPOLE_CASES])
def test_run_driver_dipole_events_match_cpu_reference(tmp_path, name, geometry, rows_, end,
                                                      integrator, extra):
    """The events run on the device state (pv_dipole_process); removals, reinjected positions
    and the counters in diagnostics.csv are the reference's, and the checkpoints carry the
    same generator text."""
    _run(REFERENCE, tmp_path / "cpu", geometry, rows_, end, integrator, extra)
    _run(DRIVER, tmp_path / "hip", geometry, rows_, end, integrator, extra)
    compare_runs(tmp_path / "cpu", tmp_path / "hip")
    last = rows(tmp_path / "hip" / "diagnostics.csv")[-1]
    assert int(last["removed_pairs"]) >= 1


def workflow(exe, root):
    """Runs every workflow scenario with exe under root; returns {name: CompletedProcess}."""
    out = {}

    def go(name, text):
        out[name] = run(exe, root, name, text)
        return out[name]

    for integ in ("rk4", "dopri5"):
        go(integ + "_common", params(integ + "_common", integrator=integ))
        go(integ + "_split", params(integ + "_split", "diagnosticsTime 0.03\ncheckpointTime 0.05\n",
                                    integ))
        src = f"{{root}}/{integ}_split/checkpoints/checkpoint_00000001.dat"
        go(integ + "_branch", params(integ + "_branch", "diagnosticsTime 0.03\n"
                                     f"checkpointTime 0.05\nrestartFile {src}\n", integ))
        go(integ + "_changed", params(integ + "_changed", "outputTime 0.02\ndiagnosticsTime 0.04\n"
                                      f"checkpointTime 0.03\nrestartFile {src}\n", integ))
        # Version 4 had one shared next-output clock and no event counter.
        common = root / f"{integ}_common" / "checkpoints" / "checkpoint_00000001.dat"
        if common.exists():
            lines = [line.replace("POINT_VORTEX_CHECKPOINT 6", "POINT_VORTEX_CHECKPOINT 4")
                     for line in common.read_text().splitlines()
                     if not line.startswith(("output_schedule ", "event_index "))]
            (root / f"{integ}_v4.dat").write_text("\n".join(lines) + "\n")
        go(integ + "_legacy", params(integ + "_legacy", f"restartFile {{root}}/{integ}_v4.dat\n",
                                     integ))
    protected = root / "protected"
    protected.mkdir(parents=True, exist_ok=True)
    (protected / "diagnostics.csv").write_text("previous results\n")
    go("protected", params("protected"))
    assert (protected / "diagnostics.csv").read_text() == "previous results\n"
    assert not (protected / "trajectory.csv").exists()
    go("protected_overwrite", params("protected", "overwriteRun true\n"))
    go("rounded", params("rounded", "endTime 0.1200000000000001\n"))
    go("tiny", params("tiny", "N 1\ntimeStep 1e-17\nendTime 1e-15\noutputTime 2e-16\n"))
    (root / "removed.dat").write_text("0 0 1\n0.001 0 -1\n")
    dip = "dipoleRemoval true\ndipoleRemovalDistance 0.01\ndiagnosticsTime 0.03\ncheckpointTime 0.05\n"
    go("removed", params("removed", "initialConditionFile {root}/removed.dat\n" + dip))
    go("removed_restart", params("removed_restart", "restartFile "
                                 "{root}/removed/checkpoints/checkpoint_00000001.dat\n" + dip))
    go("zero", params("zero", "endTime 0\n"))
    managed = ("N 2\nboundaryCondition infinite\nintegrator rk4\ntimeStep 0.007\n"
               "outputTime 0.04\nnumThreads 1\n")
    go("managed", managed + "endTime 0.08\nrunDirectory {root}/managed\n")
    go("managed_branch", managed + "endTime 0.12\nrestartFile "
       "{root}/managed/checkpoints/checkpoint_00000001.dat\nrunDirectory {root}/managed_branch\n")
    return out


@pytest.fixture(scope="module")
def workflows(tmp_path_factory):
    base = tmp_path_factory.mktemp("workflows")
    return {who: (base / who, workflow(exe, base / who))
            for who, exe in (("reference", REFERENCE), ("driver", DRIVER))}


def test_workflow_exit_status_and_messages_match(workflows):
    (rr, ref), (dr, drv) = workflows["reference"], workflows["driver"]
    assert ref.keys() == drv.keys()
    for name in ref:
        a, b = ref[name], drv[name]
        assert a.returncode == b.returncode, (name, a.stderr, b.stderr)
        if a.returncode:
            assert a.stderr.replace(str(rr), "{root}") == b.stderr.replace(str(dr), "{root}")
    assert drv["protected"].returncode == 1
    assert "already contains solver output" in drv["protected"].stderr


def test_workflow_outputs_match_the_reference(workflows):
    (rr, ref), (dr, drv) = workflows["reference"], workflows["driver"]
    for name, result in drv.items():
        if result.returncode:
            continue
        directory = "protected" if name == "protected_overwrite" else name
        compare_runs(rr / directory, dr / directory)


def test_workflow_schedules(workflows):
    """The schedule itself, on this driver's outputs (the reference's own expectations)."""
    root, res = workflows["driver"]
    for integ in ("rk4", "dopri5"):
        c = root / f"{integ}_common"
        for got in (times(c / "trajectory.csv"), times(c / "diagnostics.csv"),
                    checkpoint_times(c / "checkpoints")):
            assert approx_list(got, [0, .04, .08, .12, .13]), got
        assert str(c / "trajectory.csv") in res[f"{integ}_common"].stdout
        s = root / f"{integ}_split"
        assert approx_list(times(s / "diagnostics.csv"), [0, .03, .06, .09, .12, .13])
        assert approx_list(checkpoint_times(s / "checkpoints"), [0, .05, .1, .13])
        b = root / f"{integ}_branch"
        assert approx_list(times(b / "trajectory.csv"), [.05, .08, .12, .13])
        assert approx_list(times(b / "diagnostics.csv"), [.05, .06, .09, .12, .13])
        assert approx_list(checkpoint_times(b / "checkpoints"), [.1, .13])
        assert (b / "checkpoints" / "checkpoint_00000002.dat").exists()
        for full, resumed in zip(rows(s / "trajectory.csv")[-2:], rows(b / "trajectory.csv")[-2:]):
            for key in ("x", "y", "circulation", "u", "v"):
                assert abs(float(full[key]) - float(resumed[key])) <= 1e-12
        ch = root / f"{integ}_changed"
        assert approx_list(times(ch / "trajectory.csv"), [.05, .07, .09, .11, .13])
        assert approx_list(times(ch / "diagnostics.csv"), [.05, .09, .13])
        assert approx_list(checkpoint_times(ch / "checkpoints"), [.08, .11, .13])
        lg = root / f"{integ}_legacy"
        assert approx_list(times(lg / "trajectory.csv"), [.04, .08, .12, .13])
        assert approx_list(checkpoint_times(lg / "checkpoints"), [.08, .12, .13])
    assert (root / "protected" / "diagnostics.csv").read_text().startswith("time,frame,")
    assert approx_list(times(root / "rounded" / "trajectory.csv"),
                       [0, .04, .08, .1200000000000001])
    assert approx_list(times(root / "tiny" / "trajectory.csv"),
                       [0, 2e-16, 4e-16, 6e-16, 8e-16, 1e-15])
    assert times(root / "removed" / "trajectory.csv") == []
    assert approx_list(times(root / "removed" / "diagnostics.csv"), [0, .03, .06, .09, .12, .13])
    assert times(root / "removed_restart" / "trajectory.csv") == []
    assert approx_list(checkpoint_times(root / "removed_restart" / "checkpoints"), [.1, .13])
    z = root / "zero"
    assert times(z / "trajectory.csv") == [0] and checkpoint_times(z / "checkpoints") == [0]
    m = root / "managed"
    assert frames(m / "trajectory.csv") == [0, 1, 2] == frames(m / "diagnostics.csv")
    record = (m / "resolved_parameters.txt").read_text()
    assert "POINT_VORTEX_RUN_RECORD 1" in record and "backend HIP" in record
    assert "hip_arch" in record and "restarting 0" in record
    assert (m / "segments" / "segment_00000001" / "resolved_parameters.txt").exists()
    mb = root / "managed_branch"
    assert "restarting 1" in (mb / "resolved_parameters.txt").read_text()
    assert frames(mb / "trajectory.csv") == [1, 2, 3]


@pytest.mark.parametrize("integrator", ["rk4", "dopri5"])
def test_checkpoints_cross_between_the_executables(tmp_path, integrator):
    """The reference restarts from this driver's checkpoint and this driver from the
    reference's; each run matches the restart from the executable's own checkpoint."""
    extra = "diagnosticsTime 0.03\ncheckpointTime 0.05\n"
    for who, exe in (("ref", REFERENCE), ("drv", DRIVER)):
        r = run(exe, tmp_path / who, "src", params("src", extra, integrator))
        assert r.returncode == 0, r.stderr
    for exe_name, exe in (("ref", REFERENCE), ("drv", DRIVER)):
        for src_name in ("ref", "drv"):
            src = tmp_path / src_name / "src" / "checkpoints" / "checkpoint_00000001.dat"
            name = f"from_{src_name}"
            r = run(exe, tmp_path / exe_name, name, params(name, extra + f"restartFile {src}\n",
                                                           integrator))
            assert r.returncode == 0, r.stderr
    for exe_name in ("ref", "drv"):
        compare_runs(tmp_path / exe_name / "from_ref", tmp_path / exe_name / "from_drv")
    compare_runs(tmp_path / "ref" / "from_drv", tmp_path / "drv" / "from_ref")
