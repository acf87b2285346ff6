import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built libpvhip.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle

    if not pyoracle.ORACLE_LIB.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "all"], check=True,
                       capture_output=True)
    return pyoracle.Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle import pyoracle

    if not pyoracle.Reference.available():
        pytest.skip("reference CPU build (oracle/_ref) not present")
    return pyoracle.Reference()


def normwise(u, v, ur, vr):
    num = np.sqrt(np.sum((u - ur) ** 2 + (v - vr) ** 2))
    den = np.sqrt(np.sum(ur ** 2 + vr ** 2))
    return num / den if den > 0 else num


def per_vortex(u, v, ur, vr, floor_fraction=1e-3):
    mag = np.sqrt(ur ** 2 + vr ** 2)
    floor = floor_fraction * np.sqrt(np.mean(mag ** 2)) if len(mag) else 0.0
    err = np.sqrt((u - ur) ** 2 + (v - vr) ** 2)
    return float(np.max(err / np.maximum(mag, floor))) if len(mag) else 0.0
