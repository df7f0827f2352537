"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/*.h, its descriptor structs are laid out exactly like the
reference's (api/aoclfftz.h:86-311), and invalid descriptors are rejected with
the reference's validation rules (api/validate_problem.h:346-417) before any
device work."""
import ctypes
import os
import re
import subprocess

import pytest

import fftz

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
REF_API = "/root/reference/api"


def declared_functions():
    names = set()
    for f in os.listdir(INC):
        if not f.endswith(".h"):
            continue
        src = open(os.path.join(INC, f)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b((?:aoclfftz|fftwf?)_\w+)\s*\(", src):
            if not m.group(1).endswith("_t"):
                names.add(m.group(1))
        # FFTW declarations are generated per precision: X##name(...)
        if "FFTZ_FFTW_DECLARE" in src:
            body = src[src.index("#define FFTZ_FFTW_DECLARE"):src.index("FFTZ_FFTW_DECLARE(fftw_")]
            for m in re.finditer(r"X##(\w+)\s*(?:\(|\[)", body):
                if m.group(1) not in ("plan", "iodim", "iodim64", "r2r_kind", "write_char_func", "read_char_func"):
                    names.add("fftw_" + m.group(1))
                    names.add("fftwf_" + m.group(1))
        # legacy Fortran subroutines, generated per name prefix: F##name(...)
        if "FFTZ_FORTRAN_DECLARE" in src:
            body = src[src.index("#define FFTZ_FORTRAN_DECLARE"):src.index("FFTZ_FORTRAN_DECLARE(dfftw_")]
            prefixes = re.findall(r"^FFTZ_FORTRAN_DECLARE\((\w+),", src, flags=re.M)
            for m in re.finditer(r"F##(\w+)\s*\(", body):
                for pre in prefixes:
                    names.add(pre + m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = fftz.load()
    names = declared_functions()
    assert "aoclfftz_setup_f" in names and "aoclfftz_execute_io" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert "fftw_plan_many_dft" in names and "fftwf_execute_dft" in names and "fftw_version" in names
    assert "dfftw_plan_many_dft_" in names and "dfftwf_execute_dft_" in names and "sfftw_destroy_plan_" in names
    assert "fftw_plan_r2r_2d" in names and "fftwf_plan_guru64_r2r" in names
    assert set(fftz.EXPORTS) <= set(names)


def test_version_string():
    assert fftz.version().startswith("AOCL-FFTZ 6.0.0")


LAYOUT_PROG = r"""
#include <stdio.h>
#include <stddef.h>
#include "aoclfftz.h"
#define P(T, f) printf("%s.%s %zu %zu\n", #T, #f, offsetof(T, f), sizeof(((T*)0)->f));
int main(void) {
  printf("sizes %zu %zu %zu %zu %zu %zu\n", sizeof(aoclfftz_prob_desc_f), sizeof(aoclfftz_prob_desc_d),
         sizeof(aoclfftz_prob_desc_f_64_), sizeof(aoclfftz_prob_desc_d_64_), sizeof(aoclfftz_dim_t),
         sizeof(aoclfftz_dim_t_64_));
  P(aoclfftz_prob_desc_f, in) P(aoclfftz_prob_desc_f, out) P(aoclfftz_prob_desc_f, vec_rank)
  P(aoclfftz_prob_desc_f, dim_rank) P(aoclfftz_prob_desc_f, dims) P(aoclfftz_prob_desc_f, vecs)
  P(aoclfftz_prob_desc_f, flags) P(aoclfftz_prob_desc_f, pthr_fft) P(aoclfftz_prob_desc_f, cntrl_params)
  P(aoclfftz_prob_desc_d_64_, dims) P(aoclfftz_prob_desc_d_64_, cntrl_params)
  P(aoclfftz_flags_t, bit_reproducibility) P(aoclfftz_cntrl_params_t, measure_stats)
  printf("enum %d %d\n", (int)AOCLFFTZ_SUCCESS, (int)AOCLFFTZ_TIME_OUT);
  return 0;
}
"""


def layout(incdir, tmp_path, tag):
    src = tmp_path / f"layout_{tag}.c"
    exe = tmp_path / f"layout_{tag}"
    src.write_text(LAYOUT_PROG)
    subprocess.check_call(["gcc", "-I", incdir, str(src), "-o", str(exe)])
    return subprocess.check_output([str(exe)]).decode()


def test_struct_layout_matches_ctypes(tmp_path):
    out = layout(INC, tmp_path, "own")
    lines = dict(l.split(" ", 1) for l in out.strip().splitlines())
    sizes = [int(v) for v in lines["sizes"].split()]
    assert sizes[0] == ctypes.sizeof(fftz.ProbDesc) == sizes[1]
    assert sizes[2] == ctypes.sizeof(fftz.ProbDesc64) == sizes[3]
    assert sizes[4] == ctypes.sizeof(fftz.Dim) and sizes[5] == ctypes.sizeof(fftz.Dim64)
    assert int(lines["aoclfftz_prob_desc_f.flags"].split()[0]) == fftz.ProbDesc.flags.offset
    assert int(lines["aoclfftz_prob_desc_f.cntrl_params"].split()[0]) == fftz.ProbDesc.cntrl_params.offset
    assert lines["enum"] == "0 -5"


@pytest.mark.skipif(not os.path.isdir(REF_API), reason="reference headers not present")
def test_struct_layout_matches_reference_header(tmp_path):
    assert layout(INC, tmp_path, "own") == layout(REF_API, tmp_path, "ref")


def _setup(**kw):
    d = fftz.make_desc(**kw)
    return fftz.setup(d, "f"), d


@pytest.mark.parametrize("case", [
    dict(dims=[], vecs=[(1, 1, 1)]),                        # dim_rank 0
    dict(dims=[(16, 1, 1)], vecs=[]),                       # vec_rank 0
    dict(dims=[(0, 1, 1)], vecs=[(1, 16, 16)]),             # n = 0
    dict(dims=[(16, 0, 1)], vecs=[(1, 16, 16)]),            # in_stride 0
    dict(dims=[(16, 1, -1)], vecs=[(1, 16, 16)]),           # out_stride < 0
    dict(dims=[(16, 1, 1)], vecs=[(-1, 16, 16)]),           # negative batch
    dict(dims=[(16, 1, 1)], vecs=[(1, 0, 16)]),             # vec stride 0
    dict(dims=[(16, 1, 1)], vecs=[(1, 16, 16)], storage_order=1),
    dict(dims=[(16, 1, 1)], vecs=[(1, 16, 16)], transpose_mode=1),
    dict(dims=[(16, 1, 1)], vecs=[(1, 16, 16)], direction=2),
    dict(dims=[(16, 1, 1)], vecs=[(1, 16, 16)], fft_type=2),
    dict(dims=[(16, 1, 1)], vecs=[(1, 16, 16)], bit_repro=2),
    dict(dims=[(16, 1, 1)], vecs=[(1, 16, 16)], placement=2),
    dict(dims=[(16, 1, 1)], vecs=[(1, 16, 16)], inp=0x1000, out=0x2000, placement=0),  # in-place, in != out
    dict(dims=[(16, 1, 2)], vecs=[(1, 16, 16)], inp=0x1000, out=0x1000, placement=0),  # in-place strides differ
    dict(dims=[(16, 1, 1)], vecs=[(1, 16, 16)], inp=0x1000, out=0x1000, placement=1),  # OOP, in == out
    dict(dims=[(16, 1, 1)], vecs=[(1, 16, 16)], inp=0, out=0, placement=1),            # OOP, both NULL
])
def test_invalid_descriptors_rejected(case, capfd):
    h, _ = _setup(**case)
    assert not h
    err = capfd.readouterr().err
    assert "[ERROR]" in err


def test_null_descriptor_and_handles():
    lib = fftz.load()
    assert not lib.aoclfftz_setup_f(None)
    assert lib.aoclfftz_execute(None) == fftz.EXECUTION_FAILURE
    assert lib.aoclfftz_execute_io(None, 1, 2) == fftz.EXECUTION_FAILURE
    lib.aoclfftz_destroy(None)  # no-op, must not crash


def test_control_fields_normalised_not_rejected():
    # logger_mode / measure_stats / opt_level out of range are clamped
    # (validate_problem.h:309-338): validation passes, so any failure here
    # can only come from plan creation (no GPU in this container)
    d = fftz.make_desc(dims=[(16, 1, 1)], vecs=[(1, 16, 16)], inp=0, out=0x10)
    d.cntrl_params.logger_mode = 9
    d.cntrl_params.measure_stats = 1
    d.cntrl_params.opt_level = 17
    fftz.setup(d, "f")
    assert d.cntrl_params.logger_mode == 0
    assert d.cntrl_params.measure_stats == 0
    assert d.cntrl_params.opt_level == 3


@pytest.mark.parametrize("prog", ["legacy.f", "modern.f90"])
def test_fortran_callers_compile_and_link(prog, tmp_path):
    """The Fortran interface files (include/fftw_wrapper.f, fftw_wrapper.f03)
    compile with the image's flang and every subroutine / function the two
    caller programs use resolves in the library (running them needs a GPU:
    tests/test_gpu_fortran.py)."""
    from fortran_util import build_fortran, flang
    if flang() is None:
        pytest.skip("no Fortran compiler in this image")
    assert os.path.exists(build_fortran(prog, tmp_path))
