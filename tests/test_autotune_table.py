"""The autotuner's table format (CPU): written rules read back; malformed tables refused."""
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
