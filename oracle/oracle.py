"""CPU oracle for the GPU FFT library -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import
this module, and only to compute expected results.  It wraps two things:

* ``liboracle.so`` (oracle/fft_oracle.c): a plain-C FP64 restatement of the
  reference's transform -- recursive mixed-radix Cooley-Tukey
  (core/solvers/complex/ct_solver_dft.c:89-116), Bluestein with the reference's
  chirp and extended length (core/common/bluestein_utils.c:39-158,
  core/solvers/complex/bluestein_solver_dft.c:137-234), row-column N-D
  (core/solvers/complex/ndim_solver_dft.c:96-121), real-packed R2C recombine
  (core/kernels/non_dft/real_packed/packed_rdft_common.h:19-60), and the naive
  long-double DFT of test/bench/dft_reference.c:556-636.
* ``_ref/libaocl_fftz_ref.so`` when present: the reference library itself,
  compiled from its sources by oracle/ref.mk (see reflib.py).

Every array crossing this boundary is numpy; complex data is complex128
interleaved, strides in complex elements.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

FORWARD = -1
BACKWARD = +1


def build() -> None:
    """Compile oracle/fft_oracle.c into oracle/liboracle.so (gcc)."""
    root = os.path.dirname(_HERE)
    subprocess.check_call(["make", "-s", "-B", "-f", "oracle/Makefile"], cwd=root)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    i64, dp, ip = ctypes.c_int64, ctypes.POINTER(ctypes.c_double), ctypes.c_int
    lib.orc_dft_naive.argtypes = [i64, dp, i64, dp, i64, ip]
    lib.orc_fft1d.argtypes = [i64, dp, i64, dp, i64, ip]
    lib.orc_bluestein_length.argtypes = [i64]
    lib.orc_bluestein_length.restype = i64
    lib.orc_c2c.argtypes = [dp, dp, ip, ctypes.POINTER(i64), ip, ctypes.POINTER(i64), ip]
    lib.orc_r2c_packed_1d.argtypes = [i64, dp, dp]
    lib.orc_r2c_1d.argtypes = [i64, dp, i64, dp, i64]
    lib.orc_c2r_1d.argtypes = [i64, dp, i64, dp, i64]
    lib.orc_rel_l2.argtypes = [dp, dp, i64]
    lib.orc_rel_l2.restype = ctypes.c_double
    lib.orc_ref_rel_err.argtypes = [dp, dp, i64]
    lib.orc_ref_rel_err.restype = ctypes.c_double
    lib.orc_clear_cache.argtypes = []
    _lib = lib
    return lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c128(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.complex128))


def dft_naive(x, sign: int = FORWARD) -> np.ndarray:
    """O(N^2) DFT of each row of x (last axis), long-double accumulation."""
    lib = _load()
    x = _c128(x)
    flat = x.reshape(-1, x.shape[-1])
    out = np.empty_like(flat)
    n = flat.shape[1]
    for b in range(flat.shape[0]):
        row = np.ascontiguousarray(flat[b])
        o = np.empty_like(row)
        lib.orc_dft_naive(n, _dp(row.view(np.float64)), 1, _dp(o.view(np.float64)), 1, sign)
        out[b] = o
    return out.reshape(x.shape)


def fft(x, sign: int = FORWARD) -> np.ndarray:
    """FP64 FFT of each row of x along the last axis (any length)."""
    x = _c128(x)
    n = x.shape[-1]
    batch = int(np.prod(x.shape[:-1], dtype=np.int64)) if x.ndim > 1 else 1
    out = np.empty_like(x)
    c2c_strided(x.reshape(-1), out.reshape(-1), [(n, 1, 1)], [(batch, n, n)], sign)
    return out


def c2c_strided(xin: np.ndarray, xout: np.ndarray, dims, vecs, sign: int) -> None:
    """General batched N-D C2C on flat complex128 buffers.

    dims / vecs are lists of (n, in_stride, out_stride) triples (complex
    elements), dims[0] innermost, exactly as in the aoclfftz descriptor.
    xin and xout may be the same array (in-place)."""
    lib = _load()
    assert xin.dtype == np.complex128 and xout.dtype == np.complex128
    assert xin.flags.c_contiguous and xout.flags.c_contiguous
    d = np.ascontiguousarray(np.array(dims, dtype=np.int64).reshape(-1))
    v = np.ascontiguousarray(np.array(vecs, dtype=np.int64).reshape(-1))
    lib.orc_c2c(_dp(xin.view(np.float64)), _dp(xout.view(np.float64)), len(dims),
                d.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(vecs),
                v.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), sign)


def fftn(x, sign: int = FORWARD, batch_dims: int = 0) -> np.ndarray:
    """Dense N-D transform over the trailing (ndim - batch_dims) axes."""
    x = _c128(x)
    shape = x.shape
    tdims = shape[batch_dims:]
    nb = int(np.prod(shape[:batch_dims], dtype=np.int64)) if batch_dims else 1
    ntot = int(np.prod(tdims, dtype=np.int64))
    dims = []
    stride = 1
    for n in reversed(tdims):
        dims.append((n, stride, stride))
        stride *= n
    out = np.empty_like(x)
    c2c_strided(x.reshape(-1), out.reshape(-1), dims, [(nb, ntot, ntot)], sign)
    return out


def r2c_packed(x) -> np.ndarray:
    """Real-packed R2C (even n) of each row: restates the reference's
    packed solver (FFT of n/2 complex + recombine); returns n/2+1 bins."""
    lib = _load()
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    n = x.shape[-1]
    assert n % 2 == 0
    flat = x.reshape(-1, n)
    out = np.empty((flat.shape[0], n // 2 + 1), dtype=np.complex128)
    for b in range(flat.shape[0]):
        row = np.ascontiguousarray(flat[b])
        o = np.empty(n // 2 + 1, dtype=np.complex128)
        lib.orc_r2c_packed_1d(n, _dp(row), _dp(o.view(np.float64)))
        out[b] = o
    return out.reshape(x.shape[:-1] + (n // 2 + 1,))


def r2c(x) -> np.ndarray:
    """R2C by definition along the last axis (any n); n/2+1 bins."""
    lib = _load()
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    n = x.shape[-1]
    flat = x.reshape(-1, n)
    out = np.empty((flat.shape[0], n // 2 + 1), dtype=np.complex128)
    for b in range(flat.shape[0]):
        row = np.ascontiguousarray(flat[b])
        o = np.empty(n // 2 + 1, dtype=np.complex128)
        lib.orc_r2c_1d(n, _dp(row), 1, _dp(o.view(np.float64)), 1)
        out[b] = o
    return out.reshape(x.shape[:-1] + (n // 2 + 1,))


def c2r(z, n: int) -> np.ndarray:
    """C2R (backward, unnormalised) along the last axis; z has n/2+1 bins."""
    lib = _load()
    z = _c128(z)
    flat = z.reshape(-1, z.shape[-1])
    assert flat.shape[1] == n // 2 + 1
    out = np.empty((flat.shape[0], n), dtype=np.float64)
    for b in range(flat.shape[0]):
        row = np.ascontiguousarray(flat[b])
        o = np.empty(n, dtype=np.float64)
        lib.orc_c2r_1d(n, _dp(row.view(np.float64)), 1, _dp(o), 1)
        out[b] = o
    return out.reshape(z.shape[:-1] + (n,))


def rfftn(x, batch_dims: int = 0) -> np.ndarray:
    """N-D R2C over trailing axes: R2C on the last axis, then C2C forward on
    the others (row-column, as the reference's packed 3-D path)."""
    x = np.asarray(x, dtype=np.float64)
    z = r2c(x)
    nd = x.ndim - batch_dims
    if nd > 1:
        # C2C over the leading transform axes: move them last one at a time
        for ax in range(batch_dims, x.ndim - 1):
            zt = np.moveaxis(z, ax, -1)
            zt = fft(np.ascontiguousarray(zt), FORWARD)
            z = np.moveaxis(zt, -1, ax)
    return np.ascontiguousarray(z)


def bluestein_length(n: int) -> int:
    return int(_load().orc_bluestein_length(n))


def rel_l2(a, b) -> float:
    """||a - b||_2 / ||b||_2 with b the oracle (complex or real arrays)."""
    a = np.ascontiguousarray(np.asarray(a).astype(np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64))
    b = np.ascontiguousarray(np.asarray(b).astype(a.dtype))
    av, bv = a.reshape(-1).view(np.float64), b.reshape(-1).view(np.float64)
    return float(_load().orc_rel_l2(_dp(np.ascontiguousarray(av)), _dp(np.ascontiguousarray(bv)), av.size))


def ref_rel_err(a, b) -> float:
    """The reference's accuracy metric (test/bench/utils/compare.c:88-147)."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).reshape(-1).view(np.float64)
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128)).reshape(-1).view(np.float64)
    return float(_load().orc_ref_rel_err(_dp(a), _dp(b), a.size))


def clear_cache() -> None:
    _load().orc_clear_cache()
