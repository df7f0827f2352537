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


def test_string_forms_and_malformed_text():
    fftz.tune_clear()
    assert fftz.tune_text().startswith("#") and len(fftz.tune_text().splitlines()) == 1
    assert fftz.tune_parse(f"# c\n\n{KEY} 64 0 0 0\n") == 1
    assert lines_per_tile() == 1
    assert f"{KEY} 64 0 0 0" in fftz.tune_text()
    # a malformed line rejects the whole text and keeps the cache as it was
    assert fftz.tune_parse(f"{KEY} 128 0 0 0\nbroken line\n") == -1
    assert fftz.tune_parse(f"{KEY} 128 0 0 0 9\n") == -1
    assert fftz.tune_parse(f"{KEY} -1 0 0 0\n") == -1
    assert lines_per_tile() == 1
    fftz.tune_clear()
    assert lines_per_tile() == 4


def test_fftw_wisdom_calls_use_the_cache(tmp_path):
    """fftw[f]_export_wisdom_* / import_wisdom_* / forget_wisdom carry the tuned
    cache (the reference's versions are stubs: exports return 1 and write
    nothing, imports return 0)."""
    import ctypes
    lib = fftz.load()
    libc = ctypes.CDLL(None)
    lib.fftwf_export_wisdom_to_string.restype = ctypes.c_void_p
    lib.fftw_import_wisdom_from_string.argtypes = [ctypes.c_char_p]
    lib.fftwf_export_wisdom_to_filename.argtypes = [ctypes.c_char_p]
    lib.fftw_import_wisdom_from_filename.argtypes = [ctypes.c_char_p]
    lib.fftwf_forget_wisdom()
    assert lines_per_tile() == 4
    assert lib.fftw_import_wisdom_from_string(f"{KEY} 64 0 0 0\n".encode()) == 1
    assert lines_per_tile() == 1
    p = lib.fftwf_export_wisdom_to_string()
    assert p
    text = ctypes.string_at(p).decode()
    libc.free(ctypes.c_void_p(p))
    assert f"{KEY} 64 0 0 0" in text
    f = tmp_path / "wisdom"
    assert lib.fftwf_export_wisdom_to_filename(str(f).encode()) == 1
    assert f"{KEY} 64 0 0 0" in f.read_text()
    # callback export collects the same text
    out = []
    WF = ctypes.CFUNCTYPE(None, ctypes.c_char, ctypes.c_void_p)
    wf = WF(lambda c, d: out.append(c))
    lib.fftw_export_wisdom.argtypes = [WF, ctypes.c_void_p]
    lib.fftw_export_wisdom(wf, None)
    assert b"".join(out).decode() == text
    lib.fftw_forget_wisdom()
    assert lines_per_tile() == 4
    assert lib.fftw_import_wisdom_from_string(b"not wisdom\n") == 0
    assert lib.fftw_import_wisdom_from_filename(str(tmp_path / "missing").encode()) == 0
    assert lib.fftw_import_wisdom_from_filename(str(f).encode()) == 1
    assert lines_per_tile() == 1
    # callback import: chars until EOF
    data = iter(f"{KEY} 0 0 0 0\n".encode())
    RF = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)
    rf = RF(lambda d: next(data, -1))
    lib.fftwf_import_wisdom.argtypes = [RF, ctypes.c_void_p]
    assert lib.fftwf_import_wisdom(rf, None) == 1
    assert lines_per_tile() == 4
    assert lib.fftw_import_system_wisdom() == 0
    lib.fftw_forget_wisdom()


def test_fftw_wisdom_file_forms(tmp_path):
    """fftw_export_wisdom_to_file / fftw_import_wisdom_from_file on C FILE*
    streams opened by the caller."""
    import ctypes
    lib = fftz.load()
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    lib.fftwf_export_wisdom_to_file.argtypes = [ctypes.c_void_p]
    lib.fftw_import_wisdom_from_file.argtypes = [ctypes.c_void_p]
    lib.fftw_import_wisdom_from_string.argtypes = [ctypes.c_char_p]
    lib.fftw_forget_wisdom()
    assert lib.fftw_import_wisdom_from_string(f"{KEY} 64 0 0 0\n".encode()) == 1
    path = str(tmp_path / "w.txt").encode()
    fp = libc.fopen(path, b"w")
    assert fp
    lib.fftwf_export_wisdom_to_file(fp)
    assert libc.fclose(fp) == 0
    lib.fftw_forget_wisdom()
    assert lines_per_tile() == 4
    fp = libc.fopen(path, b"r")
    assert fp
    assert lib.fftw_import_wisdom_from_file(fp) == 1
    assert libc.fclose(fp) == 0
    assert lines_per_tile() == 1
    assert lib.fftw_import_wisdom_from_file(None) == 0
    lib.fftw_forget_wisdom()
    assert lines_per_tile() == 4


KEY24 = "p8:0:0:1:0|16777216,1,1;|1,16777216,16777216;"


def test_extended_fields_choose_factorisation():
    """The optional fields of a tuned-cache line pick the radix order of a
    pass length (r=), complex values per thread (e=) and the four-step split
    (f=); a dry-run plan of the problem shows them, and export writes them
    back unchanged."""
    fftz.tune_clear()
    pj = fftz.plan_json([(1024, 1, 1)], [(4096, 1024, 1024)], "f", 0)
    assert pj["launches"][0]["radix"] == [16, 16, 4] and pj["launches"][0]["fixed"]
    assert fftz.tune_parse(f"{KEY} 0 0 0 0 r=1024:8.8.16 e=8\n") == 1
    ln = fftz.plan_json([(1024, 1, 1)], [(4096, 1024, 1024)], "f", 0)["launches"][0]
    assert ln["radix"] == [8, 8, 16] and not ln["fixed"] and ln["E"] == 8
    assert f"{KEY} 0 0 0 0 e=8 r=1024:8.8.16" in fftz.tune_text()
    n = 1 << 24
    dflt = fftz.plan_json([(n, 1, 1)], [(1, n, n)], "d", 0, inplace=True)
    assert [l["L"] for l in dflt["launches"]] == [256, 256, 256]
    assert fftz.tune_parse(f"{KEY24} 0 0 0 0 f=128.256.512 r=256:4.4.16\n") == 1
    pj = fftz.plan_json([(n, 1, 1)], [(1, n, n)], "d", 0, inplace=True)
    assert [l["L"] for l in pj["launches"]] == [128, 256, 512]
    assert pj["launches"][1]["radix"] == [4, 4, 16]
    assert f"{KEY24} 0 0 0 0 r=256:4.4.16 f=128.256.512" in fftz.tune_text()
    fftz.tune_clear()


def test_extended_fields_malformed():
    fftz.tune_clear()
    for bad in ("r=1000:8.8.16", "r=1024:8.8.16.", "r=1024:32.32", "r=:16", "e=12", "f=1.2",
                "f=16", "x=3", "r=1024", "f=4096.4096.4096.4096"):
        assert fftz.tune_parse(f"{KEY} 0 0 0 0 {bad}\n") == -1, bad
    assert fftz.tune_text().count("\n") == 1
    # a split whose product is not the axis length is ignored by the planner
    assert fftz.tune_parse(f"{KEY24} 0 0 0 0 f=256.256.128\n") == 1
    n = 1 << 24
    pj = fftz.plan_json([(n, 1, 1)], [(1, n, n)], "d", 0, inplace=True)
    assert [l["L"] for l in pj["launches"]] == [256, 256, 256]
    fftz.tune_clear()


def test_tuned_split_geometry_in_emulator():
    """A tuned four-step split and radix orders replayed by the CPU plan
    emulator (bounds-checked gathers / scatters, FP64 line transforms) give
    the DFT of the input."""
    import numpy as np
    import plan_emulator as E
    n = 1 << 16
    key = f"p8:0:0:0:0|{n},1,1;|2,{n},{n};"
    fftz.tune_clear()
    assert fftz.tune_parse(f"{key} 0 0 0 0 f=16.64.64 r=64:4.16\n") == 1
    pj = fftz.plan_json([(n, 1, 1)], [(2, n, n)], "d", 0)
    assert [l["L"] for l in pj["launches"]] == [16, 64, 64]
    rng = np.random.default_rng(3)
    x = rng.standard_normal(2 * n) + 1j * rng.standard_normal(2 * n)
    y = np.zeros_like(x)
    E.check_launch_limits(pj)
    work = np.zeros(pj["work_bytes"] // 16, np.complex128)
    E.run(pj, {0: x.copy(), 1: y, 2: work})
    ref = np.fft.fft(x.reshape(2, n)).reshape(-1)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-13
    fftz.tune_clear()
