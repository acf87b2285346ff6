"""Builds libpvhip.so (the HIP backend) for gfx950 with hipcc, the run driver
point_vortex_hip linked to it, and the VALU issue-cost probe pv_issue_probe, in-tree.

Usage: python -m pointvortex_amd.build [--jobs N]

Every HIP translation unit is compiled with packed FP32 instructions disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``): FP32 work here is scalar VALU code.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libpvhip.so"
# Tuned launch shapes (pointvortex_amd.autotune); the library reads launch_table.txt beside it.
TUNED = PKG / "tuning" / "launch_table.txt"
TUNED_ENSEMBLE = PKG / "tuning" / "ensemble_table.txt"
INCLUDE = ROOT / "include"

ARCH = os.environ.get("PV_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast-honor-pragmas", f"-I{INCLUDE}", f"-I{CSRC}",
]

SOURCES = ["pv_kernels.hip", "pv_batch.hip", "pv_persist.hip", "pv_api.cpp", "pv_batch_api.cpp",
           "pv_noalloc.cpp"]


def _compile(src):
    obj = BUILD / (Path(src).stem + ".o")
    cmd = [HIPCC, "-x", "hip", *COMMON, "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


DRIVER_SRC = PKG / "driver" / "point_vortex_hip.cpp"
DRIVER = LIBDIR / "point_vortex_hip"
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")


def build_driver():
    """The run driver (host C++ over the C-ABI), linked to libpvhip.so beside it."""
    tmp = DRIVER.with_suffix(".tmp")
    cmd = [CXX, "-std=c++20", "-O2", "-Wall", "-Wextra", f"-I{INCLUDE}", str(DRIVER_SRC),
           f"-L{LIBDIR}", "-lpvhip", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, DRIVER)
    return DRIVER


PROBE_SRC = CSRC / "pv_issue_probe.hip"
PROBE = LIBDIR / "pv_issue_probe"


def build_probe():
    """The VALU issue-cost probe (a stand-alone HIP program, no library)."""
    tmp = PROBE.with_suffix(".tmp")
    cmd = [HIPCC, "-x", "hip", *COMMON, str(PROBE_SRC), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"probe build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, PROBE)
    return PROBE


def build(jobs=None, verbose=False):
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs or len(SOURCES)) as ex:
        results = list(ex.map(_compile, SOURCES))
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    objs = [str(o) for o, _ in results]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_driver()
    build_probe()
    for tuned, name in ((TUNED, "launch_table.txt"), (TUNED_ENSEMBLE, "ensemble_table.txt")):
        installed = LIBDIR / name
        if tuned.exists():
            shutil.copyfile(tuned, installed)
        elif installed.exists():
            installed.unlink()
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args()
    print(build(args.jobs, args.verbose))
