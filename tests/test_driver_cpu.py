"""The run driver's error paths, which end before any GPU work, against the reference's CPU
executable: every invalid parameter file, initial-condition file and corrupt checkpoint of
tests/driver_scenarios.py must stop both with the same `error: <message>` (run directories
written as {root}) and exit status 1, leaving no output behind."""
import subprocess

import pytest

from tests.driver_scenarios import (DRIVER, REFERENCE, failure_cases, params, run,
                                    write_inputs)

pytestmark = pytest.mark.skipif(not DRIVER.exists() or not REFERENCE.exists(),
                                reason="driver or reference CPU build not present")


@pytest.fixture(scope="module")
def roots(tmp_path_factory):
    base = tmp_path_factory.mktemp("driver")
    out = {}
    for who in ("reference", "driver"):
        root = base / who
        r = run(REFERENCE, root, "ok", params("ok"))
        assert r.returncode == 0, r.stderr
        write_inputs(root, root / "ok" / "checkpoints" / "checkpoint_00000001.dat")
        out[who] = root
    return out


@pytest.mark.parametrize("name,text", failure_cases(), ids=[c[0] for c in failure_cases()])
def test_driver_fails_like_the_reference(roots, name, text):
    a = run(REFERENCE, roots["reference"], name, text)
    b = run(DRIVER, roots["driver"], name, text)
    assert a.returncode == 1 and b.returncode == 1, (a.stderr, b.stderr)
    ea = a.stderr.replace(str(roots["reference"]), "{root}")
    eb = b.stderr.replace(str(roots["driver"]), "{root}")
    assert ea.startswith("error: ") and eb == ea
    assert not (roots["driver"] / "x" / "trajectory.csv").exists()


def test_driver_command_line(tmp_path):
    for args in (["a.params", "b.params"], [str(tmp_path / "missing.params")]):
        a = subprocess.run([str(REFERENCE), *args], text=True, capture_output=True, timeout=60)
        b = subprocess.run([str(DRIVER), *args], text=True, capture_output=True, timeout=60)
        assert a.returncode == b.returncode == 1
        assert a.stderr == b.stderr and a.stderr.startswith("error: ")
