"""Hardware counters of one evaluation, measured by rocprofv3 counter passes in child processes.

`bench.py` calls `evaluation_counters()` before it touches the GPU itself.  Each pass runs
``rocprofv3 --pmc <counters> -- python bench.py --only counters --n N`` as a child: the child
runs the 256 MiB calibration copy kernel once, then one RK4 step of the FP64 disk benchmark
population (four evaluations: the pair kernel and its reduction each time).  Three passes, one
per counter group (the hardware collects at most 4 TCC counters in one pass, and FETCH_SIZE
takes 3 of them):

* FETCH_SIZE, and WRITE_SIZE: KiB per dispatch, each scaled by what the same counter reads for
  the copy kernel, which moves 256 MiB each way; the evaluation's HBM bytes are the scaled
  sums over its two kernels;
* SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 and GRBM_GUI_ACTIVE for the pair kernel: the share of
  its SIMD cycles spent issuing FP64 instructions, 4 cycles per wave64 FMA/MUL/ADD and 16 per
  transcendental (the issue costs pv_issue_probe measures), over 4 SIMDs x CUs x the kernel's
  cycles (GRBM_GUI_ACTIVE summed over the 8 XCDs, so divided by 8).

Returns None for anything a pass could not measure (no rocprofv3, a failed pass), with the
reason; the benchmark line then says so.
"""
from __future__ import annotations

import csv
import os
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
COPY_KIB = 256 * 1024          # the calibration copy reads and writes 256 MiB
XCDS = 8
PASSES = {
    "fetch": ["FETCH_SIZE"],
    "write": ["WRITE_SIZE"],
    "valu": ["SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64",
             "SQ_INSTS_VALU_TRANS_F64", "GRBM_GUI_ACTIVE"],
}


def rocprofv3():
    return shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3"
                                        if Path("/opt/rocm/bin/rocprofv3").exists() else None)


def _rows(directory):
    out = []
    for f in Path(directory).rglob("*counter_collection.csv"):
        with open(f) as stream:
            out.extend(csv.DictReader(stream))
    return out


def _per_dispatch(rows):
    """{(dispatch, kernel name): {counter: value}} summed over the rows of one dispatch."""
    out = {}
    for r in rows:
        key = (int(r["Dispatch_Id"]), r["Kernel_Name"])
        d = out.setdefault(key, {})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def _kind(name):
    if "copy_kernel" in name:
        return "copy"
    if "sym_reduce_kernel" in name:
        return "reduce"
    if "sym_kernel" in name or "pair_kernel" in name:
        return "pair"
    return None


def _run_pass(counters, n, timeout):
    exe = rocprofv3()
    if exe is None:
        raise RuntimeError("rocprofv3 not found")
    with tempfile.TemporaryDirectory(prefix="pv_counters_") as d:
        cmd = [exe, "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", "pass", "--",
               sys.executable, str(ROOT / "bench.py"), "--only", "counters", "--n", str(n)]
        env = dict(os.environ, TMPDIR="/tmp")
        r = subprocess.run(cmd, cwd=str(ROOT), env=env, capture_output=True, text=True,
                           timeout=timeout)
        rows = _rows(d)
        if not rows:
            raise RuntimeError(f"rocprofv3 {' '.join(counters)}: no counter rows (status "
                               f"{r.returncode}): {r.stderr[-400:]}")
        return _per_dispatch(rows)


def _mean(values):
    return sum(values) / len(values) if values else None


def evaluation_counters(n, cu_count=256, timeout=240, log=None):
    """Counters of the FP64 disk evaluation at N = n: {"traffic_bytes": ..., "fp64_issue":
    ..., ...}; fields that could not be measured are None, with "errors" saying why."""
    out = {"traffic_bytes": None, "fp64_issue": None, "errors": []}
    got = {}
    for name, counters in PASSES.items():
        try:
            got[name] = _run_pass(counters, n, timeout)
            if log:
                log(f"[bench] counter pass {name}: {len(got[name])} dispatches")
        except Exception as e:                      # reported in the line, not fatal
            out["errors"].append(f"{name}: {e}")
    scale = {}
    for name, counter in (("fetch", "FETCH_SIZE"), ("write", "WRITE_SIZE")):
        if name not in got:
            continue
        by_kind = {}
        for (_, kernel), vals in got[name].items():
            k = _kind(kernel)
            if k and counter in vals:
                by_kind.setdefault(k, []).append(vals[counter])
        if not by_kind.get("copy") or not by_kind.get("pair"):
            out["errors"].append(f"{name}: copy or pair kernel missing from the pass")
            continue
        scale[name] = _mean(by_kind["copy"]) / COPY_KIB
        for k in ("pair", "reduce"):
            raw = _mean(by_kind.get(k, []))
            out[f"{counter.lower()}_{k}_kib"] = raw / scale[name] if raw is not None else None
        out[f"{counter.lower()}_copy_scale"] = scale[name]
    if "fetch" in scale and "write" in scale:
        kib = sum(out.get(f"{c}_{k}_kib") or 0.0 for c in ("fetch_size", "write_size")
                  for k in ("pair", "reduce"))
        out["traffic_bytes"] = int(round(kib * 1024))
    if "valu" in got:
        vals = [v for (_, kernel), v in got["valu"].items() if _kind(kernel) == "pair"]
        if vals:
            fma = _mean([v["SQ_INSTS_VALU_FMA_F64"] for v in vals])
            mul = _mean([v["SQ_INSTS_VALU_MUL_F64"] for v in vals])
            add = _mean([v["SQ_INSTS_VALU_ADD_F64"] for v in vals])
            trans = _mean([v["SQ_INSTS_VALU_TRANS_F64"] for v in vals])
            cycles = _mean([v["GRBM_GUI_ACTIVE"] for v in vals]) / XCDS
            out["fp64_issue"] = (4.0 * (fma + mul + add) + 16.0 * trans) / (4.0 * cu_count *
                                                                            cycles)
            out["fp64_wave_instructions"] = {"fma": fma, "mul": mul, "add": add, "trans": trans}
            out["pair_kernel_cycles"] = cycles
        else:
            out["errors"].append("valu: pair kernel missing from the pass")
    return out
