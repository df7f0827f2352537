"""Python binding of the MI355X FFT library's C ABI (ctypes).

The library is ``aocl-fftz_amd/lib/libaocl_fftz.so`` (built by build.py).  The
binding exposes the reference's descriptor structs (api/aoclfftz.h:86-311)
and entry points unchanged, plus a small ``Plan`` helper that drives them with
torch tensors (device buffers) or numpy arrays (host buffers, staged by the
library).  torch is only used for device memory and streams.

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence, Tuple

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libaocl_fftz.so")

SUCCESS = 0
EXECUTION_FAILURE = -1
SETUP_FAILURE = -2
INVALID_INPUT = -3
MEMORY_FAILURE = -4
TIME_OUT = -5


class Flags(ctypes.Structure):
    _fields_ = [("fft_type", ctypes.c_uint8), ("fft_direction", ctypes.c_uint8),
                ("storage_order", ctypes.c_uint8), ("fft_placement", ctypes.c_uint8),
                ("transpose_mode", ctypes.c_uint8), ("bit_reproducibility", ctypes.c_uint8)]


class Dim(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("in_stride", ctypes.c_int32), ("out_stride", ctypes.c_int32)]


class Dim64(ctypes.Structure):
    _fields_ = [("n", ctypes.c_ssize_t), ("in_stride", ctypes.c_ssize_t), ("out_stride", ctypes.c_ssize_t)]


class Smp(ctypes.Structure):
    _fields_ = [("num_threads", ctypes.c_int32), ("dynamic_load_model", ctypes.c_uint32)]


class Cntrl(ctypes.Structure):
    _fields_ = [("opt_level", ctypes.c_int32), ("opt_off", ctypes.c_int32),
                ("logger_mode", ctypes.c_int32), ("measure_stats", ctypes.c_int32)]


def _desc(dim_t):
    class ProbDesc(ctypes.Structure):
        _fields_ = [("in_", ctypes.c_void_p), ("out", ctypes.c_void_p),
                    ("vec_rank", ctypes.c_int32), ("dim_rank", ctypes.c_int32),
                    ("dims", ctypes.POINTER(dim_t)), ("vecs", ctypes.POINTER(dim_t)),
                    ("flags", Flags), ("pthr_fft", Smp), ("cntrl_params", Cntrl)]
    return ProbDesc


ProbDesc = _desc(Dim)
ProbDesc64 = _desc(Dim64)

# Every extern "C" symbol the library exports, as declared in include/*.h.
EXPORTS = (
    "aoclfftz_setup_f", "aoclfftz_setup_d", "aoclfftz_setup_f_64_", "aoclfftz_setup_d_64_",
    "aoclfftz_execute", "aoclfftz_execute_io", "aoclfftz_destroy", "aoclfftz_version",
    "aoclfftz_gpu_set_stream", "aoclfftz_gpu_synchronize", "aoclfftz_gpu_set_graph",
    "aoclfftz_gpu_set_compute_mode", "aoclfftz_gpu_autotune", "aoclfftz_gpu_tune_export",
    "aoclfftz_gpu_tune_import", "aoclfftz_gpu_tune_export_string", "aoclfftz_gpu_tune_import_string",
    "aoclfftz_gpu_tune_clear",
    "aoclfftz_gpu_server_create", "aoclfftz_gpu_server_submit", "aoclfftz_gpu_server_test",
    "aoclfftz_gpu_server_wait", "aoclfftz_gpu_server_stats", "aoclfftz_gpu_server_reset_stats",
    "aoclfftz_gpu_server_destroy",
    "aoclfftz_gpu_plan_info", "aoclfftz_gpu_plan_flops", "aoclfftz_gpu_plan_device_bytes",
    "aoclfftz_gpu_plan_json", "aoclfftz_gpu_profile", "aoclfftz_gpu_profile_read",
    "aoclfftz_gpu_plan_launches", "aoclfftz_gpu_execute_many", "aoclfftz_gpu_batch_setup",
    "aoclfftz_gpu_batch_plan_json", "aoclfftz_gpu_execute_noalloc", "aoclfftz_gpu_server_submit_timed",
)


class BatchItem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("batch", ctypes.c_int64), ("in_offset", ctypes.c_int64),
                ("out_offset", ctypes.c_int64), ("dist", ctypes.c_int64)]

_lib = None


def load():
    """Load the built library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run aocl-fftz_amd/build.py (there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    for name, dt in (("aoclfftz_setup_f", ProbDesc), ("aoclfftz_setup_d", ProbDesc),
                     ("aoclfftz_setup_f_64_", ProbDesc64), ("aoclfftz_setup_d_64_", ProbDesc64)):
        fn = getattr(lib, name)
        fn.argtypes = [ctypes.POINTER(dt)]
        fn.restype = vp
    lib.aoclfftz_execute.argtypes = [vp]
    lib.aoclfftz_execute.restype = ctypes.c_int
    lib.aoclfftz_execute_io.argtypes = [vp, vp, vp]
    lib.aoclfftz_execute_io.restype = ctypes.c_int
    lib.aoclfftz_gpu_execute_noalloc.argtypes = [vp, vp, vp]
    lib.aoclfftz_gpu_execute_noalloc.restype = ctypes.c_int
    lib.aoclfftz_destroy.argtypes = [vp]
    lib.aoclfftz_destroy.restype = None
    lib.aoclfftz_version.restype = ctypes.c_char_p
    lib.aoclfftz_gpu_set_stream.argtypes = [vp, vp]
    lib.aoclfftz_gpu_set_stream.restype = ctypes.c_int
    lib.aoclfftz_gpu_synchronize.argtypes = [vp]
    lib.aoclfftz_gpu_synchronize.restype = ctypes.c_int
    lib.aoclfftz_gpu_set_graph.argtypes = [vp, ctypes.c_int32]
    lib.aoclfftz_gpu_set_graph.restype = ctypes.c_int
    lib.aoclfftz_gpu_set_compute_mode.argtypes = [vp, ctypes.c_int32]
    lib.aoclfftz_gpu_set_compute_mode.restype = ctypes.c_int
    lib.aoclfftz_gpu_autotune.argtypes = [vp, vp, vp, ctypes.c_int32]
    lib.aoclfftz_gpu_autotune.restype = ctypes.c_int
    lib.aoclfftz_gpu_tune_export.argtypes = [ctypes.c_char_p]
    lib.aoclfftz_gpu_tune_export.restype = ctypes.c_int32
    lib.aoclfftz_gpu_tune_import.argtypes = [ctypes.c_char_p]
    lib.aoclfftz_gpu_tune_import.restype = ctypes.c_int32
    lib.aoclfftz_gpu_tune_export_string.argtypes = []
    lib.aoclfftz_gpu_tune_export_string.restype = vp  # malloc'd; released with libc free
    lib.aoclfftz_gpu_tune_import_string.argtypes = [ctypes.c_char_p]
    lib.aoclfftz_gpu_tune_import_string.restype = ctypes.c_int32
    lib.aoclfftz_gpu_tune_clear.argtypes = []
    lib.aoclfftz_gpu_tune_clear.restype = None
    lib.aoclfftz_gpu_server_create.argtypes = [ctypes.c_int64, ctypes.c_int32]
    lib.aoclfftz_gpu_server_create.restype = vp
    lib.aoclfftz_gpu_server_submit.argtypes = [vp, vp, ctypes.c_int32]
    lib.aoclfftz_gpu_server_submit.restype = ctypes.c_int64
    lib.aoclfftz_gpu_server_submit_timed.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_int32),
                                                     ctypes.POINTER(ctypes.c_double), ctypes.c_int64,
                                                     ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)]
    lib.aoclfftz_gpu_server_submit_timed.restype = ctypes.c_int64
    lib.aoclfftz_gpu_server_test.argtypes = [vp, ctypes.c_int64]
    lib.aoclfftz_gpu_server_test.restype = ctypes.c_int32
    lib.aoclfftz_gpu_server_wait.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]
    lib.aoclfftz_gpu_server_wait.restype = ctypes.c_int32
    lib.aoclfftz_gpu_server_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.c_int32]
    lib.aoclfftz_gpu_server_stats.restype = ctypes.c_int32
    lib.aoclfftz_gpu_server_reset_stats.argtypes = [vp]
    lib.aoclfftz_gpu_server_reset_stats.restype = None
    lib.aoclfftz_gpu_server_destroy.argtypes = [vp]
    lib.aoclfftz_gpu_server_destroy.restype = None
    lib.aoclfftz_gpu_plan_info.argtypes = [vp]
    lib.aoclfftz_gpu_plan_info.restype = ctypes.c_char_p
    lib.aoclfftz_gpu_plan_flops.argtypes = [vp]
    lib.aoclfftz_gpu_plan_flops.restype = ctypes.c_double
    lib.aoclfftz_gpu_plan_json.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int64]
    lib.aoclfftz_gpu_plan_json.restype = ctypes.c_int64
    lib.aoclfftz_gpu_batch_setup.argtypes = [ctypes.POINTER(BatchItem), ctypes.c_int64, ctypes.c_int32,
                                             ctypes.c_int32]
    lib.aoclfftz_gpu_batch_setup.restype = vp
    lib.aoclfftz_gpu_batch_plan_json.argtypes = [ctypes.POINTER(BatchItem), ctypes.c_int64, ctypes.c_int32,
                                                 ctypes.c_int32, ctypes.c_char_p, ctypes.c_int64]
    lib.aoclfftz_gpu_batch_plan_json.restype = ctypes.c_int64
    lib.aoclfftz_gpu_execute_many.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_int32]
    lib.aoclfftz_gpu_execute_many.restype = ctypes.c_int
    lib.aoclfftz_gpu_profile.argtypes = [vp, ctypes.c_int32]
    lib.aoclfftz_gpu_profile.restype = ctypes.c_int
    lib.aoclfftz_gpu_profile_read.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int32]
    lib.aoclfftz_gpu_profile_read.restype = ctypes.c_int32
    lib.aoclfftz_gpu_plan_launches.argtypes = [vp]
    lib.aoclfftz_gpu_plan_launches.restype = ctypes.c_int32
    lib.aoclfftz_gpu_plan_device_bytes.argtypes = [vp]
    lib.aoclfftz_gpu_plan_device_bytes.restype = ctypes.c_int64
    _lib = lib
    return lib


def _ptr(buf) -> int:
    """Address of a torch tensor, numpy array, int, or None."""
    if buf is None:
        return 0
    if isinstance(buf, int):
        return buf
    if hasattr(buf, "data_ptr"):
        return buf.data_ptr()
    if hasattr(buf, "ctypes"):
        return buf.ctypes.data
    raise TypeError(f"unsupported buffer {type(buf)}")


def make_desc(dims: Sequence[Tuple[int, int, int]], vecs: Sequence[Tuple[int, int, int]], *,
              inp=None, out=None, fft_type: int = 0, direction: int = 0, placement: Optional[int] = None,
              ilp64: bool = False, storage_order: int = 0, transpose_mode: int = 0, bit_repro: int = 0,
              threads: int = 1):
    """Build a descriptor (and keep its arrays alive on the returned object)."""
    dt, pt = (Dim64, ProbDesc64) if ilp64 else (Dim, ProbDesc)
    d = pt()
    darr = (dt * max(len(dims), 1))(*[dt(*x) for x in dims])
    varr = (dt * max(len(vecs), 1))(*[dt(*x) for x in vecs])
    d.in_ = _ptr(inp)
    d.out = _ptr(out)
    d.dim_rank = len(dims)
    d.vec_rank = len(vecs)
    d.dims = darr if dims else ctypes.POINTER(dt)()
    d.vecs = varr if vecs else ctypes.POINTER(dt)()
    if placement is None:
        placement = 0 if (_ptr(inp) == _ptr(out) and _ptr(inp) != 0) else 1
    d.flags = Flags(fft_type, direction, storage_order, placement, transpose_mode, bit_repro)
    d.pthr_fft = Smp(threads, 1)
    d.cntrl_params = Cntrl(3, 0, 0, 0)
    d._keep = (darr, varr)
    return d


def setup(desc, precision: str, ilp64: bool = False):
    lib = load()
    fn = {("f", False): lib.aoclfftz_setup_f, ("d", False): lib.aoclfftz_setup_d,
          ("f", True): lib.aoclfftz_setup_f_64_, ("d", True): lib.aoclfftz_setup_d_64_}[(precision, ilp64)]
    return fn(ctypes.byref(desc))


class Plan:
    """Owning wrapper around one library handle."""

    def __init__(self, dims, vecs, precision: str = "f", direction: int = 0, fft_type: int = 0,
                 inplace: bool = False, inp=None, out=None, ilp64: bool = False):
        self.precision = precision
        placement = 0 if inplace else 1
        if inp is None and out is None and not inplace:
            # setup with NULL buffers: out-of-place with "exactly one NULL"
            # is the documented way; use a dummy non-NULL marker for out.
            inp, out = None, 1
        self.desc = make_desc(dims, vecs, inp=inp, out=out, fft_type=fft_type, direction=direction,
                              placement=placement, ilp64=ilp64)
        self.handle = setup(self.desc, precision, ilp64)
        if not self.handle:
            raise RuntimeError("aoclfftz setup failed")

    def execute(self) -> int:
        return load().aoclfftz_execute(self.handle)

    def execute_io(self, inp, out) -> int:
        return load().aoclfftz_execute_io(self.handle, _ptr(inp), _ptr(out))

    def execute_noalloc(self, inp, out) -> int:
        """aoclfftz_gpu_execute_noalloc: device buffers, static device scratch."""
        return load().aoclfftz_gpu_execute_noalloc(self.handle, _ptr(inp), _ptr(out))

    def execute_many(self, ins, outs) -> int:
        """Execute on each (ins[i], outs[i]) pair, in order, in one call."""
        n = len(ins)
        a = (ctypes.c_void_p * n)(*[_ptr(x) for x in ins])
        b = (ctypes.c_void_p * n)(*[_ptr(x) for x in outs])
        return load().aoclfftz_gpu_execute_many(self.handle, a, b, n)

    def set_stream(self, stream) -> int:
        s = 0 if stream is None else (stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))
        return load().aoclfftz_gpu_set_stream(self.handle, s or None)

    def synchronize(self) -> int:
        return load().aoclfftz_gpu_synchronize(self.handle)

    def set_graph(self, enable: bool) -> int:
        return load().aoclfftz_gpu_set_graph(self.handle, 1 if enable else 0)

    # operand precision of complex64 radix-16 passes (aoclfftz_gpu_set_compute_mode)
    PLAIN, MFMA_F32, MFMA_BF16X3, MFMA_F16X2 = 0, 1, 2, 3

    def set_compute_mode(self, mode: int) -> int:
        return load().aoclfftz_gpu_set_compute_mode(self.handle, int(mode))

    def autotune(self, inp, out, reps: int = 5) -> int:
        """Time candidate tile shapes on device buffers (overwritten) and keep
        the fastest; the choice is cached for later plans of this problem."""
        return load().aoclfftz_gpu_autotune(self.handle, _ptr(inp), _ptr(out), int(reps))

    def profile(self, enable: bool) -> int:
        return load().aoclfftz_gpu_profile(self.handle, 1 if enable else 0)

    def profile_read(self):
        """Per-launch kernel durations (ms) since the last read, as a list of
        per-execute lists."""
        lib = load()
        cap = 1 << 16
        buf = (ctypes.c_float * cap)()
        n = lib.aoclfftz_gpu_profile_read(self.handle, buf, cap)
        per = lib.aoclfftz_gpu_plan_launches(self.handle)
        vals = list(buf[:min(n, cap)])
        return [vals[i:i + per] for i in range(0, len(vals), per)]

    def launches(self) -> int:
        """Kernel launches per execute."""
        return int(load().aoclfftz_gpu_plan_launches(self.handle))

    @property
    def info(self) -> str:
        return load().aoclfftz_gpu_plan_info(self.handle).decode()

    @property
    def flops(self) -> float:
        return load().aoclfftz_gpu_plan_flops(self.handle)

    def destroy(self) -> None:
        if self.handle:
            load().aoclfftz_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _items(items):
    arr = (BatchItem * max(1, len(items)))()
    for i, it in enumerate(items):
        n, b, io, oo = it[:4]
        arr[i] = BatchItem(n, b, io, oo, it[4] if len(it) > 4 else 0)
    return arr


class BatchPlan(Plan):
    """Uneven batch: items are (n, batch, in_offset, out_offset[, dist])."""

    def __init__(self, items, precision: str = "f", direction: int = 0):
        self.precision = precision
        self._arr = _items(items)
        self.handle = load().aoclfftz_gpu_batch_setup(self._arr, len(items), 4 if precision == "f" else 8,
                                                      direction)
        if not self.handle:
            raise RuntimeError("aoclfftz_gpu_batch_setup failed")


class Server:
    """Serving runtime (aoclfftz_gpu_server_*): submit requests from Python,
    wait for them, read latency / throughput statistics."""

    STATS = ("completed", "failed", "p50_ms", "p99_ms", "max_ms", "transforms_per_s", "gflops",
             "peak_footprint_bytes", "plans_resident", "queue_peak", "span_s")

    def __init__(self, hbm_budget_bytes: int, max_inflight: int = 8):
        self.handle = load().aoclfftz_gpu_server_create(int(hbm_budget_bytes), int(max_inflight))
        if not self.handle:
            raise RuntimeError("aoclfftz_gpu_server_create failed")

    def submit(self, dims, vecs, precision: str = "f", direction: int = 0, fft_type: int = 0,
               inp=None, out=None, inplace: bool = False) -> int:
        d = make_desc(dims, vecs, inp=inp, out=inp if inplace else out, fft_type=fft_type, direction=direction,
                      placement=0 if inplace else 1, ilp64=True)
        return load().aoclfftz_gpu_server_submit(self.handle, ctypes.byref(d), 4 if precision == "f" else 8)

    def submit_timed(self, requests, arrival_s):
        """Open-loop submission from native code: requests[i] is a dict of
        submit() keyword arguments (dims, vecs, precision, direction, fft_type,
        inp, out, inplace), submitted arrival_s[i] seconds after the call
        starts.  Returns (ids, largest lateness behind the schedule in s)."""
        n = len(requests)
        descs = []
        for r in requests:
            inplace = r.get("inplace", False)
            descs.append(make_desc(r["dims"], r["vecs"], inp=r["inp"], out=r["inp"] if inplace else r["out"],
                                   fft_type=r.get("fft_type", 0), direction=r.get("direction", 0),
                                   placement=0 if inplace else 1, ilp64=True))
        probs = (ctypes.c_void_p * max(1, n))(*[ctypes.addressof(d) for d in descs])
        precs = (ctypes.c_int32 * max(1, n))(*[4 if r.get("precision", "f") == "f" else 8 for r in requests])
        arr = (ctypes.c_double * max(1, n))(*[float(t) for t in arrival_s])
        ids = (ctypes.c_int64 * max(1, n))()
        late = ctypes.c_double(0.0)
        rc = load().aoclfftz_gpu_server_submit_timed(self.handle, probs, precs, arr, n, ids, ctypes.byref(late))
        if rc < 0:
            raise ValueError("aoclfftz_gpu_server_submit_timed rejected its arguments")
        return list(ids[:n]), late.value

    def test(self, rid: int) -> int:
        return load().aoclfftz_gpu_server_test(self.handle, rid)

    def wait(self, rid: int):
        lat = ctypes.c_double(0.0)
        rc = load().aoclfftz_gpu_server_wait(self.handle, rid, ctypes.byref(lat))
        return rc, lat.value

    def stats(self) -> dict:
        buf = (ctypes.c_double * 16)()
        n = load().aoclfftz_gpu_server_stats(self.handle, buf, 16)
        return dict(zip(self.STATS, list(buf[:n])))

    def reset_stats(self) -> None:
        load().aoclfftz_gpu_server_reset_stats(self.handle)

    def destroy(self) -> None:
        if self.handle:
            load().aoclfftz_gpu_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def batch_plan_json(items, precision: str = "f", direction: int = 0) -> dict:
    import json
    lib = load()
    arr = _items(items)
    pb = 4 if precision == "f" else 8
    n = lib.aoclfftz_gpu_batch_plan_json(arr, len(items), pb, direction, None, 0)
    if n < 0:
        raise ValueError("batch plan rejected")
    buf = ctypes.create_string_buffer(int(n) + 1)
    lib.aoclfftz_gpu_batch_plan_json(arr, len(items), pb, direction, buf, n + 1)
    return json.loads(buf.value.decode())


def plan_json(dims, vecs, precision: str = "f", direction: int = 0, fft_type: int = 0,
              inplace: bool = False) -> dict:
    """Dry-run plan (no GPU needed): the launch list as a dict."""
    import json
    lib = load()
    placement = 0 if inplace else 1
    d = make_desc(dims, vecs, inp=0x1000, out=0x1000 if inplace else 0x2000, fft_type=fft_type,
                  direction=direction, placement=placement, ilp64=True)
    pb = 4 if precision == "f" else 8
    n = lib.aoclfftz_gpu_plan_json(ctypes.byref(d), pb, 1, None, 0)
    if n < 0:
        raise ValueError("plan rejected")
    buf = ctypes.create_string_buffer(int(n) + 1)
    lib.aoclfftz_gpu_plan_json(ctypes.byref(d), pb, 1, buf, n + 1)
    return json.loads(buf.value.decode())


def tune_export(path: str) -> int:
    return load().aoclfftz_gpu_tune_export(path.encode())


def tune_import(path: str) -> int:
    return load().aoclfftz_gpu_tune_import(path.encode())


def tune_text() -> str:
    """The tuned cache as text, one problem per line."""
    p = load().aoclfftz_gpu_tune_export_string()
    if not p:
        raise MemoryError("tune cache export")
    try:
        return ctypes.string_at(p).decode()
    finally:
        ctypes.CDLL(None).free(ctypes.c_void_p(p))


def tune_parse(text: str) -> int:
    """Import cache entries from text: entries stored, -1 (nothing stored) if malformed."""
    return load().aoclfftz_gpu_tune_import_string(text.encode())


def tune_clear() -> None:
    load().aoclfftz_gpu_tune_clear()


def version() -> str:
    return load().aoclfftz_version().decode()
