"""The autotuner's cache file (cilreg_tune_save / _load / _entries / _clear).

These calls touch no device, so they run without a GPU: a file written by
hand loads, saves back to the same entries, replaces entries with the same
key, and a malformed file is rejected whole with the line named.
"""
import pytest

HEADER = "# cilreg autotune cache v1\n"
ROWS = ["gfx950:256 1 0 1 1 128 64 64 8 16 2 0.0123\n",
        "gfx950:256 0 0 1 0 64 32 32 4 0 1 0.5\n"]


@pytest.fixture
def D():
    from ccpi.filters import device
    device.tune_clear()
    yield device
    device.tune_clear()


def _entries(path):
    return sorted(l.split() for l in open(path) if l.strip() and not l.startswith("#"))


def test_load_save_round_trip(D, tmp_path):
    src = tmp_path / "a.txt"
    src.write_text(HEADER + "# a comment\n\n" + "".join(ROWS))
    assert D.tune_load(src) == 2
    assert D.tune_entries() == 2
    dst = tmp_path / "b.txt"
    assert D.tune_save(dst) == 2
    assert open(dst).readline() == HEADER
    assert [[float(v) if "." in v else v for v in e] for e in _entries(dst)] == \
           [[float(v) if "." in v else v for v in e] for e in _entries(src)]
    assert not (tmp_path / "b.txt.tmp").exists()


def test_same_key_replaces(D, tmp_path):
    a, b = tmp_path / "a.txt", tmp_path / "b.txt"
    a.write_text(HEADER + ROWS[0])
    b.write_text(HEADER + ROWS[0].replace(" 8 16 2 ", " 4 32 1 "))
    D.tune_load(a)
    D.tune_load(b)
    assert D.tune_entries() == 1
    out = tmp_path / "c.txt"
    D.tune_save(out)
    assert _entries(out)[0][8:11] == ["4", "32", "1"]


@pytest.mark.parametrize("bad, line", [
    (HEADER + ROWS[0] + "gfx950:256 1 0 1 1 128 64 64 5 16 2 0.1\n", 3),   # tile height 5
    (HEADER + "gfx950:256 9 0 1 1 128 64 64 8 16 2 0.1\n", 2),               # unknown method
    (HEADER + "gfx950:256 1 0 1 1 128 64 64 8 16 3 0.1\n", 2),               # depth 3
    (HEADER + "gfx950:256 1 0 1 1 128 64\n", 2),                             # short line
    (HEADER + ROWS[1].rstrip("\n") + " extra\n", 2),                         # trailing field
])
def test_malformed_file_rejected_whole(D, tmp_path, bad, line):
    p = tmp_path / "bad.txt"
    p.write_text(bad)
    with pytest.raises(RuntimeError, match=f"line {line}"):
        D.tune_load(p)
    assert D.tune_entries() == 0


def test_wrong_header_and_missing_file(D, tmp_path):
    p = tmp_path / "x.txt"
    p.write_text(ROWS[0])
    with pytest.raises(RuntimeError, match="not a cilreg autotune cache"):
        D.tune_load(p)
    with pytest.raises(RuntimeError, match="cannot read"):
        D.tune_load(tmp_path / "missing.txt")
    with pytest.raises(RuntimeError, match="cannot write"):
        D.tune_save(tmp_path / "no_such_dir" / "c.txt")
    assert D.tune_entries() == 0
