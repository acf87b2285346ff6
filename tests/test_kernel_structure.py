"""Compiled-structure check of the persistent ensemble kernels (CPU only: hipcc cross-compiles).

Each wave of a persistent kernel takes work items from a shared counter: lane 0 performs the
atomic add and broadcasts the item to the wave.  That is only correct when the loop the wave
re-enters for its next item is the one that contains the atomic add.  A lane-0-only branch
later in the loop body (e.g. one lane publishing a wave reduction) lets the compiler
specialise the loop header for the other lanes, producing an inner loop that re-broadcasts a
stale item without fetching a new one; the wave then never leaves it.  This test compiles
the kernels to gfx950 assembly and checks, from the compiler's loop annotations, that the
block loading each work item sits in the same outermost loop as the atomic add that fetched
it.  (The adaptive kernels failed this check before their error reduction stopped using a
lane-0-only atomic.)
"""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "pointvortex_amd" / "csrc"


def _hipcc():
    return shutil.which("hipcc") or ("/opt/rocm/bin/hipcc"
                                     if Path("/opt/rocm/bin/hipcc").exists() else None)


@pytest.fixture(scope="module")
def batch_asm(tmp_path_factory):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not available")
    from pointvortex_amd.build import COMMON

    out = tmp_path_factory.mktemp("asm") / "pv_batch.s"
    flags = [f for f in COMMON if f not in ("-fPIC",)]
    cmd = [hipcc, "-x", "hip", *flags, "--cuda-device-only", "-S", str(CSRC / "pv_batch.hip"),
           "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out.read_text()


def _kernels(asm):
    names = re.findall(r"\n(_ZN\w*batch_(?:sym_)?kernel\w*):", asm)
    for name in names:
        start = asm.index("\n" + name + ":")
        yield name, asm[start:asm.index("s_endpgm", start)].split("\n")


def _loop_of_each_line(lines):
    """(header, depth) of the innermost loop that holds each line's basic block, from the
    compiler's block annotations, and each loop header's parent; None outside every loop."""
    label = r"^(?:\.LBB\d+_\d+:|; %bb\.\d+:)"
    cur, out, parent, pending = None, [], {}, []
    for line in lines:
        m = re.match(r"^\.(LBB\d+_\d+):\s*;\s*=>\s*This (?:Inner )?Loop Header: Depth=(\d+)", line)
        if m:
            cur = (m.group(1)[1:], int(m.group(2)))
            parent[cur[0]] = None
            pending = []
            out.append(cur)
            continue
        m = re.match(r"^\.(LBB\d+_\d+):\s*;\s*Parent Loop (BB\d+_\d+) Depth=(\d+)", line)
        if m:
            # A nested loop header: its ancestors follow, then the header line itself.
            pending = [(m.group(2), int(m.group(3)))]
            cur = (m.group(1)[1:], None)
            out.append(cur)
            continue
        m = re.match(r"^\s*;\s*Parent Loop (BB\d+_\d+) Depth=(\d+)", line)
        if m and pending:
            pending.append((m.group(1), int(m.group(2))))
            out.append(cur)
            continue
        m = re.match(r"^\s*;\s*=>\s*This (?:Inner )?Loop Header: Depth=(\d+)", line)
        if m and pending:
            depth = int(m.group(1))
            parent[cur[0]] = next(h for h, d in pending if d == depth - 1)
            cur = (cur[0], depth)
            pending = []
            out[-1] = cur
            out.append(cur)
            continue
        m = re.match(label + r"\s*;\s*in Loop: Header=(BB\d+_\d+) Depth=(\d+)", line)
        if m:
            cur = (m.group(1), int(m.group(2)))
        elif re.match(label, line):
            cur = None
        out.append(cur)
    return out, parent


def test_every_persistent_kernel_was_found(batch_asm):
    names = [n for n, _ in _kernels(batch_asm)]
    # ordered: {fp64, mixed, both in one launch} x {fixed, adaptive}; symmetric: the same
    # x {per-instance widths, periodic 4-target items, disk 4-target items}
    assert len(names) == 6 + 6 * 3, names


def test_item_fetch_and_item_load_share_the_outermost_loop(batch_asm):
    """The fetch sits in the outermost loop; the item load sits in that same loop (ordered
    kernels: one item per fetch) or in a loop nested directly in it (symmetric kernels: one
    chunk of consecutive items per fetch, its bounds wave-uniform)."""
    for name, lines in _kernels(batch_asm):
        loops, parent = _loop_of_each_line(lines)
        fetch = [i for i, l in enumerate(lines) if "global_atomic_add" in l]
        assert len(fetch) == 1, (name, fetch)
        loads = [i for i, l in enumerate(lines) if i > fetch[0] and "global_load" in l]
        outer = loops[fetch[0]]
        assert outer is not None and outer[1] == 1, (name, outer)
        if "batch_sym_kernel" in name:
            # The chunk's bounds are read where the chunk was fetched; its items in the loop
            # directly inside that one.
            assert loops[loads[0]] == outer, (name, outer, loops[loads[0]])
            item = next(i for i in loads if loops[i] is not None and loops[i][1] == 2)
            assert parent[loops[item][0]] == outer[0], (name, loops[item], outer)
        else:
            assert loops[loads[0]] == outer, (name, outer, loops[loads[0]])
