"""Fortran callers of the FFTW interface, compiled with the image's flang and
linked against the library the way an existing Fortran program would be:

* tests/fortran/legacy.f -- fixed-form program calling the legacy
  subroutines (dfftw_*, dfftwf_*, sfftw_*; the reference's
  wrappers/fftw/src/fortran/*.c) with include/fftw_wrapper.f;
* tests/fortran/modern.f90 -- Fortran 2003 program through iso_c_binding and
  include/fftw_wrapper.f03 (the role of the reference's
  wrappers/fftw/api/fftw_wrapper.f03).

Inputs are complex tones and cosines whose DFT is known exactly (a peak of
height N at one bin), so a dimension-order or sign mistake moves the peak;
each program prints PASS only when every bin matches (FP64 1e-10, FP32
1e-3 absolute, with peaks of 64-192).
"""
import os

import pytest

from fortran_util import build_fortran

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prog", ["legacy.f", "modern.f90"])
def test_fortran_callers_run(prog, tmp_path):
    import subprocess
    exe = build_fortran(prog, tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS" in r.stdout and "FAIL" not in r.stdout
