"""Writes tests/golden/*.npz from the reference's own CPU code (oracle/_ref/libpvref.so).

Run in the build container, where the reference tree is present:
    make -C oracle ref && python tests/golden/make_golden.py

Each fixture holds the seeded inputs and the reference's outputs: velocities (evaluate),
invariants (computeInvariants), 10 fixed RK4 steps, a short adaptive DOPRI5 run with its
accepted-step records, and the population after one DipoleManager::process call.
reference_periodic_state.npz holds the same outputs, dipoles aside (its clusters have no
opposite-sign pair closer than 0.08), for a 200-vortex periodic state the reference keeps
among its plotting examples (`python tests/golden/make_golden.py state` writes only that
one).
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.pyoracle import Reference, spec  # noqa: E402
from pointvortex_amd import workloads as wl  # noqa: E402

OUT = Path(__file__).resolve().parent


def population(geometry, n, seed):
    if geometry == "infinite":
        return wl.infinite(n, seed)
    if geometry == "periodic":
        return wl.periodic(n, seed, box_length=2.0)
    # Near-wall vortices only where no time stepping follows (they leave the disk at dt=1e-3).
    return wl.disk(n, seed, gaussian=True, wall_fraction=0.01 if n >= 1000 else 0.0)


def main():
    ref = Reference()
    for geometry in ("infinite", "periodic", "disk"):
        data = {}
        core = 0.05 if geometry == "infinite" else 0.0
        k = spec(geometry, core_radius=0.0, box_length=2.0, image_layers=8, radius=1.0)
        kc = spec(geometry, core_radius=core, box_length=2.0, image_layers=8, radius=1.0)
        for n in (2, 4, 258, 1024):
            pop = population(geometry, n, seed=1000 + n)
            u, v = ref.evaluate(k, pop.x, pop.y, pop.circulation)
            inv = ref.invariants(kc, pop.x, pop.y, pop.circulation)
            tag = f"n{n}"
            data[f"{tag}_x"], data[f"{tag}_y"], data[f"{tag}_g"] = pop.x, pop.y, pop.circulation
            data[f"{tag}_u"], data[f"{tag}_v"] = u, v
            data[f"{tag}_invariants_core"] = inv
            data[f"{tag}_core"] = np.array([core])
            if n in (4, 258):
                dt = 1e-4 if geometry == "disk" else 1e-3
                xr, yr = ref.rk4(k, pop.x, pop.y, pop.circulation, dt, 10)
                data[f"{tag}_rk4_dt"] = np.array([dt])
                data[f"{tag}_rk4_x"], data[f"{tag}_rk4_y"] = xr, yr
                xd, yd, rec = ref.dopri5_run(k, pop.x, pop.y, pop.circulation, 1e-3, 0.01,
                                             1e-12, 1e-9, 1e-12, 0.1)
                data[f"{tag}_dopri_x"], data[f"{tag}_dopri_y"] = xd, yd
                data[f"{tag}_dopri_record"] = rec
        # Dipole removal on a population with planted close pairs (removal distance 0.03).
        pop = population(geometry, 258, seed=77)
        g = np.where(np.arange(258) % 2 == 0, 1.0, -1.0)
        x, y = pop.x.copy(), pop.y.copy()
        for a in range(0, 40, 2):
            x[a + 1] = x[a] + 0.004 * (1 + a % 5)
            y[a + 1] = y[a] - 0.003
        if geometry == "disk":
            x[100], y[100] = 0.985, 0.0
        xs, ys, gs, events = ref.dipole_process(geometry, 2.0, 1.0, 0.03, 0, 1234567, x, y, g)
        data["dipole_in_x"], data["dipole_in_y"], data["dipole_in_g"] = x, y, g
        data["dipole_out_x"], data["dipole_out_y"], data["dipole_out_g"] = xs, ys, gs
        data["dipole_events"] = np.array([events])
        np.savez_compressed(OUT / f"reference_{geometry}.npz", **data)
        print(geometry, sorted(data)[:4], "...", len(data), "arrays")


# A state the reference itself holds: 200 vortices of +-0.005 in the 2 pi periodic box, from
# its plotting examples (scripts/plotting/vortex_state/vortex_xy.006500; x y circulation per
# line).  Only its numbers are kept, with the reference's outputs for them.
STATE = Path("/root/reference/scripts/plotting/vortex_state/vortex_xy.006500")
TWO_PI = 2.0 * np.pi


def periodic_state():
    ref = Reference()
    a = np.loadtxt(STATE)
    x, y, g = (np.ascontiguousarray(a[:, c]) for c in range(3))
    k = spec("periodic", box_length=TWO_PI, image_layers=8)
    data = {"x": x, "y": y, "g": g}
    data["u"], data["v"] = ref.evaluate(k, x, y, g)
    data["invariants"] = ref.invariants(k, x, y, g)
    data["rk4_dt"] = np.array([1e-2])
    data["rk4_x"], data["rk4_y"] = ref.rk4(k, x, y, g, 1e-2, 10)
    xd, yd, rec = ref.dopri5_run(k, x, y, g, 1e-2, 0.5, 1e-12, 1e-9, 1e-12, 0.1)
    data["dopri_x"], data["dopri_y"], data["dopri_record"] = xd, yd, rec
    np.savez_compressed(OUT / "reference_periodic_state.npz", **data)
    print("periodic state", len(x), "vortices,", len(rec), "DOPRI5 steps")


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "state":
        periodic_state()
    else:
        main()
        periodic_state()
