"""The tuned-parameter cache (aoclfftz_gpu_tune_import / _export) without a
GPU: an imported entry changes the tile shape of a dry-run plan of that
problem, and export writes it back out."""
import fftz

KEY = "p4:0:0:0:0|1024,1,1;|4096,1024,1024;"


def lines_per_tile():
    pj = fftz.plan_json([(1024, 1, 1)], [(4096, 1024, 1024)], "f", 0)
    return pj["launches"][0]["C"]


def test_import_changes_plan_and_export_round_trips(tmp_path):
    assert lines_per_tile() == 4  # default: 256-thread tiles of 4 lines
    f = tmp_path / "tune.txt"
    f.write_text(f"# tuned\n{KEY} 64 0 0 0\n")
    try:
        assert fftz.tune_import(str(f)) == 1
        assert lines_per_tile() == 1  # row tile of 64 threads: one line
        out = tmp_path / "out.txt"
        assert fftz.tune_export(str(out)) >= 1
        assert f"{KEY} 64 0 0 0" in out.read_text()
    finally:
        reset = tmp_path / "reset.txt"
        reset.write_text(f"{KEY} 0 0 0 0\n")  # empty entry: planner defaults
        assert fftz.tune_import(str(reset)) == 1
    assert lines_per_tile() == 4
    assert fftz.tune_import(str(tmp_path / "missing.txt")) == -1
