"""The autotuner's table format (CPU): written rules read back; malformed tables refused."""
import pytest

from pointvortex_amd import autotune


def test_table_round_trip(tmp_path):
    results = [
        {"geometry": "disk", "precision": "fp64", "n": n, "default": (1, 8, 256),
         "default_ms": 1.0, "best": (2, 4, 128), "best_ms": 0.5}
        for n in (1024, 4096)
    ]
    path = tmp_path / "t.txt"
    path.write_text("\n".join(autotune.table_lines(results)) + "\n")
    rules = autotune.read_table(path)
    assert rules == [("disk", "fp64", 1024, 2, 4, 128),
                     ("disk", "fp64", autotune.SIZE_MAX, 2, 4, 128)]


@pytest.mark.parametrize("text", ["disk fp64 10 1 0\n", "cube fp64 10 1 0 128\n",
                                  "disk fp32 10 1 0 128\n"])
def test_malformed_tables_are_refused(tmp_path, text):
    path = tmp_path / "t.txt"
    path.write_text(text)
    with pytest.raises(ValueError):
        autotune.read_table(path)


def test_committed_table_parses():
    if autotune.TABLE.exists():
        rules = autotune.read_table(autotune.TABLE)
        assert all(r[3] in (1, 2, 4) and r[5] in (128, 256) for r in rules)
