"""ctypes binding to the reference AOCL-FFTZ library built from its sources
(oracle/ref.mk -> oracle/_ref/libaocl_fftz_ref.so).  Test infrastructure only:
used to cross-check the oracle restatement and as the timed CPU baseline.

The descriptor structs restate api/aoclfftz.h:86-311 (LP64 variants) so the
reference's own setup/execute entry points can be driven from numpy buffers.
"""
from __future__ import annotations

import ctypes
import os
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = os.path.join(_HERE, "_ref", "libaocl_fftz_ref.so")


class Flags(ctypes.Structure):
    _fields_ = [("fft_type", ctypes.c_uint8), ("fft_direction", ctypes.c_uint8),
                ("storage_order", ctypes.c_uint8), ("fft_placement", ctypes.c_uint8),
                ("transpose_mode", ctypes.c_uint8), ("bit_reproducibility", ctypes.c_uint8)]


class Dim(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("in_stride", ctypes.c_int32), ("out_stride", ctypes.c_int32)]


class Smp(ctypes.Structure):
    _fields_ = [("num_threads", ctypes.c_int32), ("dynamic_load_model", ctypes.c_uint32)]


class Cntrl(ctypes.Structure):
    _fields_ = [("opt_level", ctypes.c_int32), ("opt_off", ctypes.c_int32),
                ("logger_mode", ctypes.c_int32), ("measure_stats", ctypes.c_int32)]


class ProbDesc(ctypes.Structure):
    _fields_ = [("in_", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("vec_rank", ctypes.c_int32), ("dim_rank", ctypes.c_int32),
                ("dims", ctypes.POINTER(Dim)), ("vecs", ctypes.POINTER(Dim)),
                ("flags", Flags), ("pthr_fft", Smp), ("cntrl_params", Cntrl)]


def available() -> bool:
    return os.path.exists(REF_LIB)


_lib = None


def _load():
    global _lib
    if _lib is None:
        # RTLD_LOCAL (ctypes default) keeps these symbols from colliding with
        # the GPU library's identically named entry points.
        lib = ctypes.CDLL(REF_LIB)
        for name in ("aoclfftz_setup_f", "aoclfftz_setup_d"):
            getattr(lib, name).argtypes = [ctypes.POINTER(ProbDesc)]
            getattr(lib, name).restype = ctypes.c_void_p
        lib.aoclfftz_execute.argtypes = [ctypes.c_void_p]
        lib.aoclfftz_execute.restype = ctypes.c_int
        lib.aoclfftz_execute_io.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.aoclfftz_execute_io.restype = ctypes.c_int
        lib.aoclfftz_destroy.argtypes = [ctypes.c_void_p]
        lib.aoclfftz_version.restype = ctypes.c_char_p
        _lib = lib
    return _lib


class RefPlan:
    """One reference plan over caller-owned numpy buffers."""

    def __init__(self, dims, vecs, precision: str, direction: int, inp: np.ndarray,
                 out: np.ndarray, threads: int = 1, fft_type: int = 0, opt_level: int = 3):
        lib = _load()
        self._dims = (Dim * len(dims))(*[Dim(*d) for d in dims])
        self._vecs = (Dim * len(vecs))(*[Dim(*v) for v in vecs])
        pd = ProbDesc()
        pd.in_ = inp.ctypes.data
        pd.out = out.ctypes.data
        pd.vec_rank = len(vecs)
        pd.dim_rank = len(dims)
        pd.dims = self._dims
        pd.vecs = self._vecs
        pd.flags = Flags(fft_type, 0 if direction < 0 else 1, 0,
                         0 if inp.ctypes.data == out.ctypes.data else 1, 0, 0)
        pd.pthr_fft = Smp(threads, 1)
        pd.cntrl_params = Cntrl(opt_level, 0, 0, 0)
        self._pd = pd
        self._keep = (inp, out)
        setup = lib.aoclfftz_setup_f if precision == "f" else lib.aoclfftz_setup_d
        self.handle = setup(ctypes.byref(pd))
        if not self.handle:
            raise RuntimeError("reference setup failed")

    def execute(self) -> int:
        return _load().aoclfftz_execute(self.handle)

    def execute_io(self, inp: np.ndarray, out: np.ndarray) -> int:
        return _load().aoclfftz_execute_io(self.handle, inp.ctypes.data, out.ctypes.data)

    def destroy(self) -> None:
        if self.handle:
            _load().aoclfftz_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def c2c(x: np.ndarray, direction: int = -1, precision: str = "d", threads: int = 1) -> np.ndarray:
    """Batched 1-D C2C of each row of x (last axis) through the reference."""
    dt = np.complex64 if precision == "f" else np.complex128
    x = np.ascontiguousarray(np.asarray(x, dtype=dt))
    n = x.shape[-1]
    b = x.size // n
    out = np.empty_like(x)
    p = RefPlan([(n, 1, 1)], [(b, n, n)], precision, direction, x, out, threads)
    rc = p.execute()
    p.destroy()
    if rc != 0:
        raise RuntimeError("reference execute failed")
    return out


def time_c2c(n: int, batch: int, precision: str, threads: int, budget_s: float = 10.0,
             inplace: bool = False, seed: int = 1):
    """Time the reference's batched C2C forward transform on host memory.

    Calibrated like test/bench/performance.c:76-106 (warm-up, then repeats until
    the budget is used); returns (min seconds per execute, repeats)."""
    dt = np.complex64 if precision == "f" else np.complex128
    rng = np.random.default_rng(seed)
    x = (rng.uniform(-1, 1, (batch, n)) + 1j * rng.uniform(-1, 1, (batch, n))).astype(dt)
    out = x if inplace else np.empty_like(x)
    p = RefPlan([(n, 1, 1)], [(batch, n, n)], precision, -1, x, out, threads)
    p.execute()
    p.execute()
    best = float("inf")
    reps = 0
    t_end = time.perf_counter() + budget_s
    while True:
        t0 = time.perf_counter()
        p.execute()
        dt_s = time.perf_counter() - t0
        best = min(best, dt_s)
        reps += 1
        if time.perf_counter() > t_end and reps >= 3:
            break
    p.destroy()
    return best, reps
