"""GPU parity through the FFTW3-compatible interface (boundary #2).

Covers the reference wrapper's translation rules
(wrappers/fftw/src/translator/fftz_translator.c:379-600, :712-806): plan_many
with embed arrays and strides, guru dims, r2c / c2r defaults, in-place, and
execute_dft on fresh arrays; plus two C programs compiled and linked the way
an existing caller would (tests/c/*.c)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import fftz
from gpu_util import tol
from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("prog", ["example_aoclfftz", "example_fftw"])
def test_c_callers_link_and_run(prog, tmp_path):
    exe = tmp_path / prog
    lib = os.path.join(ROOT, "aocl-fftz_amd", "lib")
    subprocess.check_call(["gcc", "-O2", "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "c", prog + ".c"),
                           "-o", str(exe), "-L" + lib, "-laocl_fftz", "-lm", "-Wl,-rpath," + lib])
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "FAIL" not in r.stdout and "PASS" in r.stdout


def _lib():
    lib = fftz.load()
    vp, i32, ip = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)
    for pre in ("fftw_", "fftwf_"):
        f = getattr(lib, pre + "plan_many_dft")
        f.argtypes = [i32, ip, i32, vp, ip, i32, i32, vp, ip, i32, i32, i32, ctypes.c_uint]
        f.restype = vp
        for k in ("plan_many_dft_r2c", "plan_many_dft_c2r"):
            g = getattr(lib, pre + k)
            g.argtypes = [i32, ip, i32, vp, ip, i32, i32, vp, ip, i32, i32, ctypes.c_uint]
            g.restype = vp
        getattr(lib, pre + "execute").argtypes = [vp]
        getattr(lib, pre + "execute_dft").argtypes = [vp, vp, vp]
        getattr(lib, pre + "destroy_plan").argtypes = [vp]
        getattr(lib, pre + "plan_guru64_dft").argtypes = [i32, vp, i32, vp, vp, vp, i32, ctypes.c_uint]
        getattr(lib, pre + "plan_guru64_dft").restype = vp
    return lib


def iarr(v):
    return None if v is None else (ctypes.c_int * len(v))(*v)


def rnd(shape, seed, dt):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(dt)


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("case", [
    # rank, n, howmany, inembed, istride, idist, onembed, ostride, odist
    (1, [64], 5, None, 1, 64, None, 1, 64),
    (1, [100], 3, None, 3, 1, None, 1, 100),
    (2, [12, 20], 2, [14, 22], 1, 14 * 22, None, 1, 240),
    (3, [4, 6, 8], 2, None, 1, 192, [5, 6, 9], 2, 600),
    (1, [101], 2, None, 1, 101, None, 1, 101),
])
@pytest.mark.parametrize("sign", [-1, 1])
def test_plan_many_dft(prec, case, sign):
    lib = _lib()
    rank, n, hm, ie, ist, idist, oe, ost, odist = case
    dt = np.complex64 if prec == "f" else np.complex128
    iemb = ie or n
    oemb = oe or n
    in_len = (hm - 1) * idist + ist * int(np.prod(iemb)) + 1
    out_len = (hm - 1) * odist + ost * int(np.prod(oemb)) + 1
    x = rnd(in_len, 5, dt)
    y = np.zeros(out_len, dt)
    pre = "fftwf_" if prec == "f" else "fftw_"
    p = getattr(lib, pre + "plan_many_dft")(rank, iarr(n), hm, x.ctypes.data, iarr(ie), ist, idist,
                                             y.ctypes.data, iarr(oe), ost, odist, sign, 0)
    assert p
    getattr(lib, pre + "execute")(p)
    getattr(lib, pre + "destroy_plan")(p)
    # oracle through the equivalent aoclfftz descriptor (dims reversed)
    dims = []
    si, so = ist, ost
    for i in range(rank):
        dims.append((n[rank - 1 - i], si, so))
        si *= iemb[rank - 1 - i]
        so *= oemb[rank - 1 - i]
    ref = np.zeros(out_len, np.complex128)
    O.c2c_strided(x.astype(np.complex128), ref, dims, [(hm, idist, odist)], sign)
    assert O.rel_l2(y, ref) <= tol(int(np.prod(n)), prec)


def test_execute_dft_new_arrays_and_inplace():
    lib = _lib()
    n, hm = 512, 4
    x = rnd(n * hm, 1, np.complex64)
    y = np.zeros_like(x)
    p = lib.fftwf_plan_many_dft(1, iarr([n]), hm, x.ctypes.data, None, 1, n, y.ctypes.data, None, 1, n, -1, 0)
    x2 = rnd(n * hm, 2, np.complex64)
    y2 = np.zeros_like(x2)
    lib.fftwf_execute_dft(p, x2.ctypes.data, y2.ctypes.data)
    lib.fftwf_destroy_plan(p)
    assert O.rel_l2(y2, O.fft(x2.reshape(hm, n).astype(np.complex128)).reshape(-1)) <= tol(n, "f")
    # in-place plan (in == out)
    z = x.copy()
    p = lib.fftwf_plan_many_dft(1, iarr([n]), hm, z.ctypes.data, None, 1, n, z.ctypes.data, None, 1, n, 1, 0)
    lib.fftwf_execute(p)
    lib.fftwf_destroy_plan(p)
    assert O.rel_l2(z, O.fft(x.reshape(hm, n).astype(np.complex128), +1).reshape(-1)) <= tol(n, "f")


def test_guru64_dims():
    lib = _lib()

    class IOD(ctypes.Structure):
        _fields_ = [("n", ctypes.c_ssize_t), ("is", ctypes.c_ssize_t), ("os", ctypes.c_ssize_t)]
    # 2-D 8 x 10 transform (row-major), batch of 3 with distance 100 / 80
    dims = (IOD * 2)(IOD(8, 10, 10), IOD(10, 1, 1))
    hd = (IOD * 1)(IOD(3, 100, 80))
    x = rnd(300, 4, np.complex128)
    y = np.zeros(240, np.complex128)
    p = lib.fftw_plan_guru64_dft(2, ctypes.byref(dims), 1, ctypes.byref(hd), x.ctypes.data, y.ctypes.data, -1, 0)
    assert p
    lib.fftw_execute(p)
    lib.fftw_destroy_plan(p)
    for b in range(3):
        ref = np.fft.fft2(x[b * 100:b * 100 + 80].reshape(8, 10))
        assert O.rel_l2(y[b * 80:(b + 1) * 80], ref.reshape(-1)) <= tol(80, "d")


# odd and prime innermost lengths run the full-length real path
@pytest.mark.parametrize("shape", [(64,), (16, 20), (8, 6, 12), (63,), (10, 101), (9, 7, 5)])
def test_r2c_c2r_defaults(shape):
    lib = _lib()
    rank = len(shape)
    h = shape[-1] // 2 + 1
    hm = 2
    nreal = int(np.prod(shape))
    ncplx = int(np.prod(shape[:-1])) * h
    rng = np.random.default_rng(9)
    xr = rng.standard_normal(nreal * hm)
    y = np.zeros(ncplx * hm, np.complex128)
    p = lib.fftw_plan_many_dft_r2c(rank, iarr(list(shape)), hm, xr.ctypes.data, None, 1, nreal, y.ctypes.data,
                                   None, 1, ncplx, 0)
    assert p
    lib.fftw_execute(p)
    lib.fftw_destroy_plan(p)
    ref = np.fft.rfftn(xr.reshape((hm,) + shape), axes=tuple(range(1, rank + 1)))
    assert O.rel_l2(y, ref.reshape(-1)) <= tol(nreal, "d")
    back = np.zeros(nreal * hm)
    yc = y.copy()
    p = lib.fftw_plan_many_dft_c2r(rank, iarr(list(shape)), hm, yc.ctypes.data, None, 1, ncplx, back.ctypes.data,
                                   None, 1, nreal, 0)
    assert p
    lib.fftw_execute(p)
    lib.fftw_destroy_plan(p)
    assert O.rel_l2(back / nreal, xr) <= 2 * tol(nreal, "d")
    # the reference keeps the Hermitian input of an out-of-place C2R intact
    assert np.array_equal(yc, y)
