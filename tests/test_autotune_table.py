"""The autotuner's table formats (CPU): written rules read back; malformed tables refused; the
committed tables parse.  Single systems (launch_table.txt) and ensembles (ensemble_table.txt)."""
import pytest

from pointvortex_amd import autotune


def test_table_round_trip(tmp_path):
    results = [
        {"geometry": "disk", "precision": "fp64", "n": n, "default": (1, 8, 256, 0),
         "default_ms": 1.0, "best": (2, 4, 128, sym), "best_ms": 0.5}
        for n, sym in ((1024, 0), (4096, 1))
    ]
    path = tmp_path / "t.txt"
    path.write_text("\n".join(autotune.table_lines(results)) + "\n")
    rules = autotune.read_table(path)
    assert rules == [("disk", "fp64", 1024, 2, 4, 128, 0),
                     ("disk", "fp64", autotune.SIZE_MAX, 2, 4, 128, 1)]


def test_six_column_rules_mean_split_path(tmp_path):
    path = tmp_path / "t.txt"
    path.write_text("infinite mixed 100 1 2 128\n")
    assert autotune.read_table(path) == [("infinite", "mixed", 100, 1, 2, 128, 0)]


def test_partial_slot_count():
    assert autotune.partial_slots((2, 8, 256, 0), 65536) == 8
    assert autotune.partial_slots((2, 1, 256, 0), 65536) == 0
    assert autotune.partial_slots((4, 0, 256, 1), 65536) == 64


@pytest.mark.parametrize("text", ["disk fp64 10 1 0\n", "cube fp64 10 1 0 128\n",
                                  "disk fp32 10 1 0 128\n", "disk fp64 10 1 0 128 2\n"])
def test_malformed_tables_are_refused(tmp_path, text):
    path = tmp_path / "t.txt"
    path.write_text(text)
    with pytest.raises(ValueError):
        autotune.read_table(path)


def test_committed_table_parses():
    if autotune.TABLE.exists():
        rules = autotune.read_table(autotune.TABLE)
        assert all(r[3] in (1, 2, 4, 8) and r[5] in (128, 256) and r[6] in (0, 1) for r in rules)


def test_partial_slots_of_eight_targets_per_thread():
    # 128 threads x 8 = 1024-vortex tiles, as at 4 targets per thread
    assert autotune.partial_slots((8, 0, 256, 1), 65536) == 64


def test_ensemble_table_round_trip(tmp_path):
    results = [
        {"geometry": "periodic", "precision": "mixed", "n": n, "count": 16,
         "default": (2, 4), "default_ms": 2.0, "best": best, "best_ms": 1.5}
        for n, best in ((256, (1, 8)), (4096, (4, 2)))
    ]
    path = tmp_path / "e.txt"
    path.write_text("\n".join(autotune.ensemble_table_lines(results)) + "\n")
    assert autotune.read_ensemble_table(path) == [
        ("periodic", "mixed", 256, 1, 8), ("periodic", "mixed", autotune.SIZE_MAX, 4, 2)]


@pytest.mark.parametrize("text", ["disk fp64 10 1\n", "cube fp64 10 1 4\n",
                                  "disk fp32 10 1 4\n", "disk fp64 10 1 4 0\n"])
def test_malformed_ensemble_tables_are_refused(tmp_path, text):
    path = tmp_path / "e.txt"
    path.write_text(text)
    with pytest.raises(ValueError):
        autotune.read_ensemble_table(path)


def test_ensemble_instance_counts():
    assert autotune.ensemble_count(64) == 4096
    assert autotune.ensemble_count(1024) == 256
    assert autotune.ensemble_count(65536) == 4


def test_committed_ensemble_table_parses():
    if autotune.ENSEMBLE_TABLE.exists():
        rules = autotune.read_ensemble_table(autotune.ENSEMBLE_TABLE)
        assert rules and all(r[3] in (1, 2, 4) and 1 <= r[4] <= 64 for r in rules)


def test_ensemble_mix_choice_selects_rules(tmp_path):
    """Per precision the table keeps the shapes its mix check chose: the per-bucket ones, the
    shared kernel's, or none (the built-in shape)."""
    results = [
        {"geometry": "disk", "precision": prec, "n": 4096, "count": 16, "default": (2, 4),
         "default_ms": 2.0, "best": (4, 4), "best_ms": 1.5, "best_shared": (2, 8)}
        for prec in ("fp64", "mixed")
    ]
    mix = {"fp64": ("built-in", {"best": 3.0, "best_shared": 2.5, "built-in": 2.0}),
           "mixed": ("best_shared", {"best": 3.0, "best_shared": 1.0, "built-in": 2.0})}
    path = tmp_path / "e.txt"
    path.write_text("\n".join(autotune.ensemble_table_lines(results, mix)) + "\n")
    assert autotune.read_ensemble_table(path) == [("disk", "mixed", autotune.SIZE_MAX, 2, 8)]
