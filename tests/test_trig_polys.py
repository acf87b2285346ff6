"""The periodic kernels' trigonometric polynomials (pointvortex_amd/csrc/pv_device.hpp).

The helpers cos2_* (2 cos h) and sinh_sin2_* (2 sinh h and 2 sin h) carry fitted coefficients
as hex literals.  This test reads those literals back from the header and checks the
polynomials against mpmath over their whole range |h| <= pi/2 (a mistyped digit or a
coefficient in the wrong place fails it), at the error bounds their comments state.
"""
import re
from pathlib import Path

import numpy as np
import pytest

mp = pytest.importorskip("mpmath")

HEADER = Path(__file__).resolve().parents[1] / "pointvortex_amd" / "csrc" / "pv_device.hpp"
HEX = r"-?0x[0-9a-f]+(?:\.[0-9a-f]*)?p[+-]?\d+f?"


def body(name):
    text = HEADER.read_text()
    start = text.index(name + "(")
    return text[start:text.index("\n}\n", start)]


def chains(name):
    """Horner chains of a helper: variable -> coefficients, highest power first."""
    out = {}
    for line in body(name).splitlines()[1:]:
        m = re.match(r"\s*(?:(?:const )?(?:double|float) )?(\w+) = .*?(" + HEX + r")\)?;", line)
        r = re.match(r"\s*return .*?(" + HEX + r")\);", line)
        if m and m.group(1) in ("a", "b", "p"):
            out.setdefault(m.group(1), []).append(m.group(2))
        elif r:
            out["p"].append(r.group(1))
    return {k: [float.fromhex(v.rstrip("f")) for v in vals] for k, vals in out.items()}


def horner(coefs, z):
    r = mp.mpf(0)
    for c in coefs:
        r = r * z + mp.mpf(c)
    return r


Z = (mp.pi / 2) ** 2
GRID = 1501


@pytest.mark.parametrize("name,degree,bound", [("cos2_f64", 9, 4e-17), ("cos2_f32", 5, 8e-8)])
def test_cos2(name, degree, bound):
    mp.mp.dps = 40
    c = chains(name)["p"]
    assert len(c) == degree + 1
    if name.endswith("f32"):
        assert all(np.float32(x) == x for x in c)
    err = max(abs(horner(c, z) - 2 * mp.cos(mp.sqrt(z))) for z in mp.linspace(0, Z, GRID))
    assert err <= bound


@pytest.mark.parametrize("name,degree,bound", [("sinh_sin2_f64", 9, 6e-17),
                                               ("sinh_sin2_f32", 5, 4e-8)])
def test_sinh_sin2(name, degree, bound):
    mp.mp.dps = 40
    ch = chains(name)
    a, b = ch["a"], ch["b"]
    assert len(a) + len(b) == degree + 1
    worst = mp.mpf(0)
    for h in mp.linspace(mp.mpf("1e-3"), mp.pi / 2, GRID):
        z = h * h
        w = z * z
        A, B = horner(a, w), horner(b, w)
        sh, sn = h * (A + z * B), h * (A - z * B)
        worst = max(worst, abs(sh / (2 * mp.sinh(h)) - 1), abs(sn / (2 * mp.sin(h)) - 1))
    assert worst <= bound
