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
