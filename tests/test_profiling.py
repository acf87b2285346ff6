"""The counter passes' arithmetic (CPU): copy-kernel calibration of FETCH_SIZE / WRITE_SIZE,
the evaluation's HBM bytes, and the FP64 issue share, on synthetic rocprofv3 rows."""
import csv

import pytest

from pointvortex_amd import profiling

FIELDS = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def write_rows(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(FIELDS, r)))


def test_rows_are_summed_per_dispatch(tmp_path):
    write_rows(tmp_path / "pass_counter_collection.csv",
               [(1, "copy_kernel", "FETCH_SIZE", 10.0), (1, "copy_kernel", "FETCH_SIZE", 5.0),
                (2, "void pv::sym_kernel<pv::DiskF64, 2, 128>", "FETCH_SIZE", 7.0)])
    got = profiling._per_dispatch(profiling._rows(tmp_path))
    assert got[(1, "copy_kernel")]["FETCH_SIZE"] == 15.0
    assert profiling._kind("void pv::sym_kernel<pv::DiskF64, 2, 128>") == "pair"
    assert profiling._kind("pv::sym_reduce_kernel(pv::SymLaunch, pv::GeoParams)") == "reduce"
    assert profiling._kind("pv::reset_status_kernel") is None


def test_traffic_and_fp64_issue(monkeypatch):
    copy_kib = profiling.COPY_KIB
    passes = {
        ("FETCH_SIZE",): {(1, "copy_kernel"): {"FETCH_SIZE": copy_kib / 2},
                          (2, "sym_kernel<DiskF64>"): {"FETCH_SIZE": 100.0},
                          (3, "sym_reduce_kernel"): {"FETCH_SIZE": 50.0}},
        ("WRITE_SIZE",): {(1, "copy_kernel"): {"WRITE_SIZE": copy_kib},
                          (2, "sym_kernel<DiskF64>"): {"WRITE_SIZE": 400.0},
                          (3, "sym_reduce_kernel"): {"WRITE_SIZE": 8.0}},
    }
    valu = {"SQ_INSTS_VALU_FMA_F64": 600.0, "SQ_INSTS_VALU_MUL_F64": 300.0,
            "SQ_INSTS_VALU_ADD_F64": 100.0, "SQ_INSTS_VALU_TRANS_F64": 50.0,
            "GRBM_GUI_ACTIVE": 8 * 10.0}

    def fake(counters, n, timeout):
        if counters[0].startswith("SQ_"):
            return {(2, "sym_kernel<DiskF64>"): valu, (3, "sym_reduce_kernel"): {}}
        return passes[tuple(counters)]

    monkeypatch.setattr(profiling, "_run_pass", fake)
    out = profiling.evaluation_counters(65536, cu_count=2)
    # FETCH scaled by 2 (the copy read half its bytes), WRITE by 1.
    assert out["fetch_size_pair_kib"] == pytest.approx(200.0)
    assert out["fetch_size_reduce_kib"] == pytest.approx(100.0)
    assert out["traffic_bytes"] == round((200.0 + 100.0 + 400.0 + 8.0) * 1024)
    # (4 (600 + 300 + 100) + 16 * 50) / (4 SIMDs x 2 CUs x 10 cycles)
    assert out["fp64_issue"] == pytest.approx((4 * 1000 + 16 * 50) / 80.0)
    assert out["errors"] == []


def test_a_failed_pass_leaves_its_fields_empty(monkeypatch):
    def fail(counters, n, timeout):
        raise RuntimeError("rocprofv3 not found")

    monkeypatch.setattr(profiling, "_run_pass", fail)
    out = profiling.evaluation_counters(65536)
    assert out["traffic_bytes"] is None and out["fp64_issue"] is None
    assert len(out["errors"]) == 3
