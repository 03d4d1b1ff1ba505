"""The autotuner's host side: it reads the shapes csrc/tuned.hpp holds, searches only values the
kernels accept, and rewrites the header in place (the GPU side, timing and the byte-identity check
of every variant's results, runs with `python -m brook_amd.tune`)."""
import shutil

from brook_amd import tune


def test_header_round_trip(tmp_path, monkeypatch):
    values = tune.current()
    assert set(values) == set(tune.SPACE)
    for name, value in values.items():
        assert value in tune.SPACE[name]
    copy = tmp_path / "tuned.hpp"
    shutil.copy(tune.HEADER, copy)
    monkeypatch.setattr(tune, "HEADER", str(copy))
    changed = dict(values, BROOK_NB_BATCH=26, BROOK_FLOOD_BLOCK=512)
    tune.write_header(changed, {"plain_ms": 1.0})
    tune.write_header(changed, {"plain_ms": 2.0})  # the note is replaced, not stacked
    assert tune.current() == changed
    text = copy.read_text()
    assert text.count("// Last tuned:") == 1 and '"plain_ms": 2.0' in text
    assert text.count("#ifndef") == len(tune.SPACE)  # -D on the compiler line still wins


def test_search_space_is_what_the_kernels_accept():
    for name in ("BROOK_NB_BATCH", "BROOK_FLOOD_NB_BATCH"):  # the chunkings for_parts implements
        assert set(tune.SPACE[name]) <= {9, 13, 26}
    for name in ("BROOK_TRACE_BLOCK", "BROOK_FLOOD_BLOCK"):
        assert all(v % 64 == 0 and v <= 1024 for v in tune.SPACE[name])  # whole waves, one workgroup
    for name in ("BROOK_TRACE_GRID", "BROOK_FLOOD_GRID"):
        assert all(v >= 0 for v in tune.SPACE[name])  # 0: two per compute unit
