"""Scenarios for the run driver (pointvortex_amd/lib/point_vortex_hip), each run by it and by
the reference's CPU executable (oracle/_ref/point_vortex_cpu, built from the reference
sources) in a root directory of its own, and the comparison of what the two leave behind.

A scenario is (name, params text, must succeed).  `{root}` in the params text is the
executable's own root, so that restarts and inputs refer to the files that run produced.  The
behaviours follow the reference's output and restart workflows (its tests/output_integration.py:
schedules, branch and changed-interval restarts, legacy v4 checkpoints, overwrite protection,
roundoff end times, a dipole-emptied population, frame numbering, run records).
"""
import csv
import math
import re
import subprocess
from pathlib import Path

from oracle import pyoracle

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "pointvortex_amd" / "lib" / "point_vortex_hip"
REFERENCE = pyoracle.REF_CPU_EXE

BASE = ("N 2\nboundaryCondition infinite\nintegrator {integrator}\ntimeStep 0.007\nnumThreads 1\n"
        "endTime 0.13\noutputTime 0.04\nrunDirectory {root}/{name}\n")


def params(name, extra="", integrator="rk4"):
    return BASE.replace("{integrator}", integrator).replace("{name}", name) + extra


def run(exe, root, name, text, timeout=120):
    """Writes {root}/{name}.params (with {root} filled in) and runs exe on it."""
    root.mkdir(parents=True, exist_ok=True)
    f = root / (name.replace("/", "_") + ".params")
    f.write_text(text.replace("{root}", str(root)))
    return subprocess.run([str(exe), str(f)], text=True, capture_output=True, timeout=timeout)


def rows(path):
    with open(path) as stream:
        return list(csv.DictReader(stream))


def times(path):
    return list(dict.fromkeys(float(r["time"]) for r in rows(path)))


def frames(path):
    return list(dict.fromkeys(int(r["frame"]) for r in rows(path)))


def checkpoint_times(directory):
    return [float(p.read_text().splitlines()[1].split()[1])
            for p in sorted(Path(directory).glob("checkpoint_*.dat"))]


def close(a, b, rel=2e-10, abs_=2e-11):
    return math.isclose(a, b, rel_tol=rel, abs_tol=abs_)


def compare_csv(a, b, rel=2e-10, abs_=2e-11):
    """Same header, same rows; integers exactly, every other field within rel/abs."""
    with open(a) as fa, open(b) as fb:
        ha, hb = fa.readline(), fb.readline()
    assert ha == hb, (a, ha, hb)
    ra, rb = rows(a), rows(b)
    assert len(ra) == len(rb), (a, len(ra), len(rb))
    for x, y in zip(ra, rb):
        for key in x:
            if key in ("frame", "index", "removed_pairs", "reinjected_pairs"):
                assert x[key] == y[key], (a, key, x[key], y[key])
            else:
                assert close(float(x[key]), float(y[key]), rel, abs_), (a, key, x[key], y[key])


def compare_checkpoint(a, b, rel=2e-10, abs_=2e-11):
    """Same fields in the same order; integers and words exactly, reals within rel/abs, the
    generator state text exactly."""
    la, lb = Path(a).read_text().splitlines(), Path(b).read_text().splitlines()
    assert len(la) == len(lb), (a, len(la), len(lb))
    for x, y in zip(la, lb):
        if x.startswith("random_engine_state"):
            assert x == y, a
            continue
        ta, tb = x.split(), y.split()
        assert len(ta) == len(tb), (a, x, y)
        for p, q in zip(ta, tb):
            if re.fullmatch(r"-?\d+", p) or not re.fullmatch(r"[-+0-9.eEinfa]+", p):
                assert p == q, (a, x, y)
            else:
                assert close(float(p), float(q), rel, abs_), (a, x, y)


def compare_runs(da, db, rel=2e-10, abs_=2e-11):
    """Two run directories: the CSVs, the checkpoint list and contents, the run records' keys."""
    for f in ("trajectory.csv", "diagnostics.csv"):
        compare_csv(da / f, db / f, rel, abs_)
    ca = sorted(p.name for p in (da / "checkpoints").glob("checkpoint_*.dat"))
    cb = sorted(p.name for p in (db / "checkpoints").glob("checkpoint_*.dat"))
    assert ca == cb, (da, ca, cb)
    for name in ca:
        compare_checkpoint(da / "checkpoints" / name, db / "checkpoints" / name, rel, abs_)
    keys = [[line.split()[0] for line in (d / "resolved_parameters.txt").read_text().splitlines()
             if line and not line.startswith(("hip_", "omp_", "openmp", "mpi_", "cuda_"))]
            for d in (da, db)]
    assert keys[0] == keys[1], keys


def failure_cases():
    """Parameter, input and checkpoint errors: both executables stop before any output with
    `error: <message>` and exit status 1.  Checkpoint cases corrupt one field of a valid
    checkpoint the reference wrote ({root}/ok/checkpoints/checkpoint_00000001.dat)."""
    cases = []
    for key in ("diagnosticsTime", "checkpointTime"):
        for value in ("0", "-1", "nan", "inf"):
            cases.append((f"{key}_{value}", params("x", f"{key} {value}\n")))
    cases += [
        ("unknown_key", params("x", "outputFile obsolete.csv\n")),
        ("missing_value", params("x", "endTime\n")),
        ("extra_value", params("x", "endTime 0.1 0.2\n")),
        ("bad_integrator", params("x", "integrator euler\n")),
        ("bad_flag", params("x", "dipoleRemoval maybe\n")),
        ("bad_reinjection", params("x", "dipoleReinjection sometimes\n")),
        ("reinjection_without_removal", params("x", "dipoleReinjection paired\n")),
        ("reinjection_infinite", params("x", "dipoleRemoval true\ndipoleReinjection paired\n")),
        ("rectangular_box", params("x", "boundaryCondition periodic\nboxLengthY 3\n")),
        ("core_radius_disk", params("x", "boundaryCondition disk\ncoreRadius 0.1\n")),
        ("too_many_layers", params("x", "periodicImageLayers 65\n")),
        ("negative_N", params("x", "N -3\n")),
        ("zero_N", params("x", "N 0\n")),
        ("bad_double", params("x", "timeStep fast\n")),
        ("bad_int", params("x", "numThreads 1.5\n")),
        ("negative_threads", params("x", "numThreads -1\n")),
        ("bad_tolerances", params("x", "absoluteTolerance 0\n")),
        ("bad_bounds", params("x", "minimumTimeStep 1\nmaximumTimeStep 0.5\n")),
        ("odd_periodic", params("x", "N 3\nboundaryCondition periodic\n")),
        ("missing_initial", params("x", "initialConditionFile {root}/nowhere.dat\n")),
        ("malformed_initial", params("x", "initialConditionFile {root}/malformed.dat\n")),
        ("empty_initial", params("x", "initialConditionFile {root}/empty.dat\n")),
        ("geometry_mismatch", params("x", "initialConditionFile {root}/tagged.dat\n")),
        ("radius_mismatch", params("x", "boundaryCondition disk\ndiskRadius 2\n"
                                        "initialConditionFile {root}/tagged_disk.dat\n")),
        ("missing_restart", params("x", "restartFile {root}/none.dat\n")),
        ("restart_mismatch", params("x", "boundaryCondition disk\n"
                                         "restartFile {root}/ok/checkpoints/checkpoint_00000001.dat\n")),
    ]
    for field, value in [("output_index", "-1"), ("event_index", "-1"),
                         ("vortex_count", "999999999999999999"), ("next_output_time", "0"),
                         ("accepted_steps", "-1"), ("time", "-1"), ("suggested_time_step", "0"),
                         ("integrator", "euler"), ("dipole_counts", "1 2"),
                         ("POINT_VORTEX_CHECKPOINT", "9")]:
        cases.append((f"corrupt_{field}",
                      params("x", f"restartFile {{root}}/corrupt_{field}.dat\n")))
    cases.append(("trailing_data", params("x", "restartFile {root}/corrupt_trailing.dat\n")))
    return cases


def write_inputs(root, reference_checkpoint):
    """The input files the failure cases read, under root."""
    root.mkdir(parents=True, exist_ok=True)
    (root / "malformed.dat").write_text("not a vortex\n0 0 1\n")
    (root / "empty.dat").write_text("# nothing here\n")
    (root / "tagged.dat").write_text("# geometry=periodic box_length=2\n0.1 0 1\n-0.1 0 -1\n")
    (root / "tagged_disk.dat").write_text("# geometry=disk disk_radius=1\n0.1 0 1\n-0.1 0 -1\n")
    lines = Path(reference_checkpoint).read_text().splitlines()
    for field, value in [("output_index", "-1"), ("event_index", "-1"),
                         ("vortex_count", "999999999999999999"), ("next_output_time", "0"),
                         ("accepted_steps", "-1"), ("time", "-1"), ("suggested_time_step", "0"),
                         ("integrator", "euler"), ("dipole_counts", "1 2"),
                         ("POINT_VORTEX_CHECKPOINT", "9")]:
        out = [f"{field} {value}" if line.startswith(field + " ") else line for line in lines]
        (root / f"corrupt_{field}.dat").write_text("\n".join(out) + "\n")
    (root / "corrupt_trailing.dat").write_text("\n".join(lines) + "\n1 2 3\n")
