"""Build helper for the Fortran caller programs under tests/fortran."""
import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "aocl-fftz_amd", "lib")


def flang():
    """The image's Fortran compiler (ROCm's flang), or None."""
    for c in (shutil.which("flang"), shutil.which("amdflang"), "/opt/rocm/llvm/bin/flang",
              "/opt/rocm/lib/llvm/bin/flang"):
        if c and os.path.exists(c):
            return c
    return None


def build_fortran(prog, outdir):
    """Compile tests/fortran/<prog> and link it against the library; returns
    the executable path."""
    fc = flang()
    if fc is None:
        raise RuntimeError("no Fortran compiler (flang) found")
    src = os.path.join(ROOT, "tests", "fortran", prog)
    exe = os.path.join(str(outdir), prog.split(".")[0])
    subprocess.check_call([fc, "-O1", "-I", INC, src, "-o", exe, "-L" + LIB, "-laocl_fftz",
                           "-Wl,-rpath," + LIB], cwd=str(outdir))
    return exe
