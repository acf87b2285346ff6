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


def disk_longdouble(x, y, g, radius=1.0):
    """Disk velocities (src/compute.cpp:301-333 formula) in x87 extended precision, as a
    truth reference for cases where the FP64 reference itself loses digits near the wall."""
    L = np.longdouble
    x, y, g = (np.asarray(a, dtype=L) for a in (x, y, g))
    R = L(radius)
    inv2pi = L(0.5) / L(np.pi)
    tx, ty = x[:, None] / R, y[:, None] / R
    sx, sy = x[None, :] / R, y[None, :] / R
    dx, dy = x[:, None] - x[None, :], y[:, None] - y[None, :]
    r2 = dx * dx + dy * dy
    np.fill_diagonal(r2, L(1))
    c = inv2pi * g[None, :] / r2
    np.fill_diagonal(c, L(0))
    a = L(1) - (tx * sx + ty * sy)
    b = ty * sx - tx * sy
    k = -inv2pi * g[None, :] / R / (a * a + b * b)
    u = (-c * dy).sum(1) - (k * (-a * sy + b * sx)).sum(1)
    v = (c * dx).sum(1) + (k * (-a * sx - b * sy)).sum(1)
    return u.astype(np.float64), v.astype(np.float64)


def disk_longdouble_targets(x, y, g, targets, radius=1.0, chunk=64):
    """disk_longdouble for the listed targets only (all sources), in chunks of targets."""
    targets = np.asarray(targets)
    L = np.longdouble
    xs, ys, gs = (np.asarray(a, dtype=L) for a in (x, y, g))
    R = L(radius)
    inv2pi = L(0.5) / L(np.pi)
    sx, sy = xs[None, :] / R, ys[None, :] / R
    u = np.zeros(len(targets))
    v = np.zeros(len(targets))
    for c0 in range(0, len(targets), chunk):
        idx = targets[c0:c0 + chunk]
        tx, ty = xs[idx, None] / R, ys[idx, None] / R
        dx, dy = xs[idx, None] - xs[None, :], ys[idx, None] - ys[None, :]
        r2 = dx * dx + dy * dy
        self_mask = np.arange(len(xs))[None, :] == idx[:, None]
        r2[self_mask] = L(1)
        c = inv2pi * gs[None, :] / r2
        c[self_mask] = L(0)
        a = L(1) - (tx * sx + ty * sy)
        b = ty * sx - tx * sy
        k = -inv2pi * gs[None, :] / R / (a * a + b * b)
        u[c0:c0 + len(idx)] = ((-c * dy).sum(1) - (k * (-a * sy + b * sx)).sum(1)).astype(np.float64)
        v[c0:c0 + len(idx)] = ((c * dx).sum(1) + (k * (-a * sx - b * sy)).sum(1)).astype(np.float64)
    return u, v


def disk_truth(x, y, g, radius=1.0):
    """Disk velocities of every target in extended precision (cached for the last state)."""
    key = (np.asarray(x).tobytes(), np.asarray(y).tobytes(), np.asarray(g).tobytes(), radius)
    if key not in _TRUTH:
        _TRUTH.clear()
        _TRUTH[key] = disk_longdouble_targets(x, y, g, np.arange(len(x)), radius)
    return _TRUTH[key]


def fp64_gate(geometry, x, y, g, ur, vr, radius=1.0):
    """Normwise tolerance for FP64 parity against the oracle: 1e-12, or twice the oracle's own
    distance from an extended-precision sum when that is larger (a vortex within ~1e-4 of the
    disk wall makes 1 - t.s cancel in the reference's FP64 arithmetic).  The HIP disk kernels
    do not share that cancellation (the image offset is formed from R^2 - |x|^2 to a few ulps),
    so the disk tests also hold them to 1e-12 of the extended-precision sum itself."""
    if geometry != "disk" or len(x) > 8192:
        return 1e-12
    ue, ve = disk_truth(x, y, g, radius)
    return max(1e-12, 2.0 * normwise(ur, vr, ue, ve))


_TRUTH = {}


def hamiltonian_longdouble(geometry, x, y, g, core_radius=0.0, box_length=2.0, layers=8,
                           radius=1.0, chunk=256):
    """The reference's Hamiltonians (src/compute.cpp:142-151, 174-202, 262-285) in x87 extended
    precision, and the rounding scale S = sum over the terms of |G_i G_j| (|e_ij| + 1) / 4 pi
    (disk self terms included).  An FP64 evaluation whose every term is accurate to a few ulps
    (of the term and of unit-size parts of e_ij) and whose sum is compensated differs from H by
    at most a small multiple of 2^-53 S."""
    L = np.longdouble
    xs, ys, gs = (np.asarray(a, dtype=L) for a in (x, y, g))
    n = len(xs)
    pi = L(np.pi) + L(1.2246467991473532e-16)       # pi to extended precision
    H = L(0)
    S = L(0)
    if geometry == "disk":
        R = L(radius)
        r2 = xs * xs + ys * ys
        self_e = np.log((R * R - r2) / R)
        H += np.sum(gs * gs * self_e) / (4 * pi)
        S += np.sum(gs * gs * (np.abs(self_e) + 1)) / (4 * pi)
    for c0 in range(0, n, chunk):
        i = np.arange(c0, min(n, c0 + chunk))
        j = np.arange(n)
        mask = j[None, :] > i[:, None]
        if not mask.any():
            continue
        dx = xs[i, None] - xs[None, :]
        dy = ys[i, None] - ys[None, :]
        if geometry == "infinite":
            r2 = dx * dx + dy * dy + L(core_radius) ** 2
            e = np.log(np.where(mask, r2, L(1)))
        elif geometry == "disk":
            R2 = L(radius) ** 2
            num = (dx * dx + dy * dy) * R2
            dot = xs[i, None] * xs[None, :] + ys[i, None] * ys[None, :]
            cross = ys[i, None] * xs[None, :] - xs[i, None] * ys[None, :]
            den = (R2 - dot) ** 2 + cross ** 2
            num = np.where(mask, num, L(1))
            e = np.log(num / den)
        else:
            Lb = L(box_length)
            k = 2 * pi / Lb
            X = k * (dx - Lb * np.rint(dx / Lb))
            Y = k * (dy - Lb * np.rint(dy / Lb))
            e = -X * X / (2 * pi)
            sy2 = np.sin(Y / 2) ** 2
            for m in range(-layers, layers + 1):
                sh = np.sinh((X - 2 * pi * m) / 2)
                d = 2 * (sh * sh + sy2)
                d = np.where(mask, d, L(1))
                a = abs(2 * pi * m)
                e = e + np.log(d) - (a + np.log1p(np.exp(-2 * a)) - np.log(L(2)))
        gg = gs[i, None] * gs[None, :]
        e = np.where(mask, e, L(0))
        H -= np.sum(gg * e) / (4 * pi)
        S += np.sum(np.where(mask, np.abs(gg) * (np.abs(e) + 1), L(0))) / (4 * pi)
    return float(H), float(S)
