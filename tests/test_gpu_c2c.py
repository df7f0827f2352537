"""GPU parity: complex-to-complex transforms through the C ABI vs the oracle.

Mirrors the reference's accuracy mode (test/bench/accuracy.c:707-855 with
ACCURACY_WITH_DFT) and its API suites (test/gtest/api/api_gtest.cpp): every
result is compared with the FP64 oracle on the same seeded input, with the
relative-L2 bound written in gpu_util.tol().
"""
import numpy as np
import pytest

import fftz
from gpu_util import rand_c, run_c2c, run_desc, to_dev, tol
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def check(y, x, prec, direction=0):
    ref = O.fft(x.astype(np.complex128), -1 if direction == 0 else 1)
    err = O.rel_l2(y, ref)
    n = x.shape[-1]
    assert err <= tol(n, prec), f"rel-L2 {err:.3e} > {tol(n, prec):.3e} (n={n})"
    return err


def test_cfg1_full_size(torch_gpu):
    """N=1024, batch 4096, complex64, out-of-place, U[-1,1) from seed 1."""
    x = rand_c((4096, 1024), 1, "f")
    y, info = run_c2c(torch_gpu, x)
    err = check(y, x, "f")
    print(f"cfg1 rel-L2 {err:.3e}  {info}")
    assert O.ref_rel_err(y, O.fft(x.astype(np.complex128))) < 1e-5


POW2 = [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384]


@pytest.mark.parametrize("n", POW2)
@pytest.mark.parametrize("prec", ["f", "d"])
def test_pow2(torch_gpu, n, prec):
    b = max(1, min(64, (1 << 16) // n))
    x = rand_c((b, n), n, prec)
    y, _ = run_c2c(torch_gpu, x)
    check(y, x, prec)


SMOOTH = [3, 5, 6, 7, 9, 10, 11, 12, 13, 15, 20, 48, 60, 100, 120, 210, 243, 343, 625, 1000, 1331,
          2187, 3003, 4096 * 3, 6561, 15625]


@pytest.mark.parametrize("n", SMOOTH)
@pytest.mark.parametrize("prec", ["f", "d"])
def test_smooth(torch_gpu, n, prec):
    b = max(1, min(16, (1 << 15) // n))
    x = rand_c((b, n), n + 7, prec)
    y, _ = run_c2c(torch_gpu, x)
    check(y, x, prec)


PRIMES = [17, 19, 23, 31, 101, 199, 1009, 4099, 8191]


@pytest.mark.parametrize("n", PRIMES)
@pytest.mark.parametrize("prec", ["f", "d"])
def test_bluestein(torch_gpu, n, prec):
    # FP64 4099 / 8191 (extended length 8232 / 16384) run the multi-pass form
    x = rand_c((3, n), n, prec)
    y, info = run_c2c(torch_gpu, x)
    assert "bluestein" in info.lower()
    check(y, x, prec)


@pytest.mark.parametrize("n", [64, 1000, 1024, 101])
@pytest.mark.parametrize("prec", ["f", "d"])
def test_backward_and_roundtrip(torch_gpu, n, prec):
    x = rand_c((4, n), 99, prec)
    y, _ = run_c2c(torch_gpu, x, direction=1)
    check(y, x, prec, direction=1)
    X, _ = run_c2c(torch_gpu, x, direction=0)
    z, _ = run_c2c(torch_gpu, X, direction=1)
    assert O.rel_l2(z / n, x) <= 2 * tol(n, prec)


@pytest.mark.parametrize("n", [256, 1024, 1000, 101])
def test_inplace(torch_gpu, n):
    x = rand_c((8, n), 5, "f")
    y, _ = run_c2c(torch_gpu, x, inplace=True)
    check(y, x, "f")


def test_strided_and_column_batches(torch_gpu):
    # 12-point transforms, element stride = batch count (column-major batch),
    # batch stride 1: the reference's column layout (data_layout_conventions.rst)
    n, b = 12, 37
    x = rand_c(n * b, 3, "f")
    out = run_desc(torch_gpu, x, n * b, [(n, b, b)], [(b, 1, 1)], "f", 0)
    ref = np.zeros(n * b, np.complex128)
    O.c2c_strided(x.astype(np.complex128), ref, [(n, b, b)], [(b, 1, 1)], -1)
    assert O.rel_l2(out, ref) <= tol(n, "f")
    # element stride 2, distance 3n, output packed
    n, b = 256, 5
    x = rand_c(2 * 3 * n * b, 4, "d")
    out = run_desc(torch_gpu, x, n * b, [(n, 2, 1)], [(b, 3 * n, n)], "d", 0)
    ref = np.zeros(n * b, np.complex128)
    O.c2c_strided(x, ref, [(n, 2, 1)], [(b, 3 * n, n)], -1)
    assert O.rel_l2(out, ref) <= tol(n, "d")


@pytest.mark.parametrize("shape", [(16, 16), (64, 48), (8, 12, 10), (32, 32, 32), (101, 8), (3, 5, 7, 4)])
@pytest.mark.parametrize("prec", ["f", "d"])
def test_multidim(torch_gpu, shape, prec):
    b = 2
    x = rand_c((b,) + shape, 17, prec)
    dims = []
    stride = 1
    for n in reversed(shape):
        dims.append((n, stride, stride))
        stride *= n
    total = stride
    out = run_desc(torch_gpu, x.reshape(-1), total * b, dims, [(b, total, total)], prec, 0)
    ref = O.fftn(x, batch_dims=1)
    assert O.rel_l2(out, ref.reshape(-1)) <= tol(total, prec)


def test_multi_vec_batch(torch_gpu):
    # two batch axes that do not fuse: 3 groups x 4 transforms, gap between groups
    n = 64
    x = rand_c(3 * 5 * n, 21, "f")
    vecs = [(4, n, n), (3, 5 * n, 4 * n)]
    out = run_desc(torch_gpu, x, 3 * 4 * n, [(n, 1, 1)], vecs, "f", 0)
    ref = np.zeros(3 * 4 * n, np.complex128)
    O.c2c_strided(x.astype(np.complex128), ref, [(n, 1, 1)], vecs, -1)
    assert O.rel_l2(out, ref) <= tol(n, "f")


def test_host_buffers_staged(torch_gpu):
    n, b = 1000, 7
    x = rand_c((b, n), 8, "d")
    y = np.zeros_like(x)
    p = fftz.Plan([(n, 1, 1)], [(b, n, n)], "d", 0, inp=x, out=y)
    assert p.execute() == 0
    check(y, x, "d")
    # execute_io on fresh host buffers, same handle
    x2 = rand_c((b, n), 9, "d")
    y2 = np.zeros_like(x2)
    assert p.execute_io(x2, y2) == 0
    check(y2, x2, "d")
    p.destroy()


def test_null_setup_buffers_execute_io(torch_gpu):
    n, b = 512, 4
    x = rand_c((b, n), 10, "f")
    p = fftz.Plan([(n, 1, 1)], [(b, n, n)], "f", 0)  # in NULL at setup
    assert p.execute() == fftz.EXECUTION_FAILURE
    dx = to_dev(torch_gpu, x)
    dy = torch_gpu.empty_like(dx)
    assert p.execute_io(dx, dy) == 0
    torch_gpu.cuda.synchronize()
    check(dy.cpu().numpy(), x, "f")
    p.destroy()


def test_zero_batch_is_noop(torch_gpu):
    dx = torch_gpu.zeros(64, dtype=torch_gpu.complex64, device="cuda")
    dy = torch_gpu.full((64,), 7.0, dtype=torch_gpu.complex64, device="cuda")
    p = fftz.Plan([(64, 1, 1)], [(0, 64, 64)], "f", 0, inp=dx, out=dy)
    assert p.execute() == 0
    torch_gpu.cuda.synchronize()
    assert (dy.cpu().numpy() == 7.0).all()
    p.destroy()


def test_size_one(torch_gpu):
    x = rand_c((5, 1), 2, "f")
    y, _ = run_c2c(torch_gpu, x)
    assert np.array_equal(y, x)
    y, _ = run_c2c(torch_gpu, x, direction=1)
    assert np.array_equal(y, x)


@pytest.mark.parametrize("n,prec", [(1024, "f"), (4096, "d"), (101, "f"), (1 << 20, "f")])
def test_bit_reproducible(torch_gpu, n, prec):
    """Same input, same plan, repeated runs: bitwise identical output
    (test/test_bit_repro.c compares value casts; this compares bytes)."""
    b = max(1, (1 << 20) // n)
    x = rand_c((b, n), 1234, prec)
    dx = to_dev(torch_gpu, x)
    dy = torch_gpu.empty_like(dx)
    p = fftz.Plan([(n, 1, 1)], [(b, n, n)], prec, 0, inp=dx, out=dy)
    outs = []
    for _ in range(3):
        assert p.execute() == 0
        torch_gpu.cuda.synchronize()
        outs.append(dy.cpu().numpy().tobytes())
    p.destroy()
    assert outs[0] == outs[1] == outs[2]


def test_graph_replay_matches_eager(torch_gpu):
    n, b = 1024, 256
    x = rand_c((b, n), 77, "f")
    eager, _ = run_c2c(torch_gpu, x)
    dx = to_dev(torch_gpu, x)
    dy = torch_gpu.empty_like(dx)
    p = fftz.Plan([(n, 1, 1)], [(b, n, n)], "f", 0, inp=dx, out=dy)
    p.set_graph(True)
    for _ in range(4):
        dy.zero_()
        assert p.execute() == 0
        torch_gpu.cuda.synchronize()
        assert dy.cpu().numpy().tobytes() == eager.tobytes()
    p.destroy()


def test_concurrent_execute_io_matches_serial(torch_gpu):
    """Concurrent execute_io on one shared handle gives exactly the serial
    results (test/gtest/api/api_gtest.h:651-786, AoclfftzConcurrentTest)."""
    import threading
    cases = [(1024, 64, "f"), (101, 10, "f"), (1 << 15, 2, "f")]
    for n, b, prec in cases:
        p = fftz.Plan([(n, 1, 1)], [(b, n, n)], prec, 0)
        xs = [to_dev(torch_gpu, rand_c((b, n), 500 + i, prec)) for i in range(6)]
        serial = []
        for xd in xs:
            yd = torch_gpu.empty_like(xd)
            assert p.execute_io(xd, yd) == 0
            serial.append(yd.cpu().numpy().tobytes())
        outs = [torch_gpu.empty_like(xd) for xd in xs]
        rcs = [None] * len(xs)

        def work(i):
            rcs[i] = p.execute_io(xs[i], outs[i])

        th = [threading.Thread(target=work, args=(i,)) for i in range(len(xs))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        torch_gpu.cuda.synchronize()
        assert all(rc == 0 for rc in rcs)
        for i in range(len(xs)):
            assert outs[i].cpu().numpy().tobytes() == serial[i]
        p.destroy()


@pytest.mark.parametrize("n", [1 << 15, 1 << 17, 1 << 20, 3 * (1 << 14), 2 * 3 * 5 * 7 * 11 * 13])
@pytest.mark.parametrize("prec", ["f", "d"])
def test_large_fourstep(torch_gpu, n, prec):
    b = max(1, (1 << 21) // n)
    x = rand_c((b, n), n, prec)
    y, info = run_c2c(torch_gpu, x)
    assert "four-step" in info
    check(y, x, prec)


def test_autotune_keeps_results(torch_gpu, tmp_path):
    """Autotuning picks a tile shape by timing candidates on the caller's
    buffers; the tuned plan's result matches the oracle and the choice lands
    in the exported cache."""
    torch = torch_gpu
    for n, b in ((1024, 512), (1 << 20, 2)):
        x = rand_c((b, n), 9, "f")
        dx = to_dev(torch, x)
        dy = torch.empty_like(dx)
        scratch = torch.empty_like(dx)
        p = fftz.Plan([(n, 1, 1)], [(b, n, n)], "f", 0, inp=dx, out=dy)
        assert p.autotune(dx.clone(), scratch, reps=3) == 0
        assert p.execute() == 0
        torch.cuda.synchronize()
        assert O.rel_l2(dy.cpu().numpy(), O.fft(x.astype(np.complex128)).reshape(-1)) <= tol(n, "f")
        print(f"n={n}: {p.info[:160]}")
        p.destroy()
    path = str(tmp_path / "tune.txt")
    assert fftz.tune_export(path) >= 2


def test_wisdom_round_trip_restores_tuned_plan(torch_gpu):
    """The tuned cache survives export -> forget -> import as FFTW wisdom text:
    a plan made after the import has the tuned tile shape and gives the tuned
    plan's result bit for bit.  A tile shape imported by hand (one line per
    64-thread tile) changes only the grouping of lines, so the result is the
    same bits as the default shape's."""
    import ctypes
    torch = torch_gpu
    lib = fftz.load()
    lib.fftw_export_wisdom_to_string.restype = ctypes.c_void_p
    lib.fftw_import_wisdom_from_string.argtypes = [ctypes.c_char_p]
    n, b = 1024, 512
    key = f"p4:0:0:0:0|{n},1,1;|{b},{n},{n};"
    x = rand_c((b, n), 19, "f")
    dx = to_dev(torch, x)

    def run():
        dy = torch.empty_like(dx)
        p = fftz.Plan([(n, 1, 1)], [(b, n, n)], "f", 0, inp=dx, out=dy)
        assert p.execute() == 0
        torch.cuda.synchronize()
        info = p.info
        p.destroy()
        return info, dy

    lib.fftw_forget_wisdom()
    info_default, y_default = run()
    scratch_in, scratch_out = dx.clone(), torch.empty_like(dx)
    p = fftz.Plan([(n, 1, 1)], [(b, n, n)], "f", 0, inp=dx, out=torch.empty_like(dx))
    assert p.autotune(scratch_in, scratch_out, reps=3) == 0
    p.destroy()
    info_tuned, y_tuned = run()
    s = lib.fftw_export_wisdom_to_string()
    text = ctypes.string_at(s)
    ctypes.CDLL(None).free(ctypes.c_void_p(s))
    assert key.encode() in text
    lib.fftw_forget_wisdom()
    assert run()[0] == info_default
    assert lib.fftw_import_wisdom_from_string(text) == 1
    info_again, y_again = run()
    assert info_again == info_tuned
    assert torch.equal(y_again.view(torch.float32), y_tuned.view(torch.float32))
    assert lib.fftw_import_wisdom_from_string(f"{key} 64 0 0 0\n".encode()) == 1
    info_64, y_64 = run()
    assert "C=1 " in info_64, info_64
    assert torch.equal(y_64.view(torch.float32), y_default.view(torch.float32))
    assert O.rel_l2(y_64.cpu().numpy(), O.fft(x.astype(np.complex128)).reshape(-1)) <= tol(n, "f")
    lib.fftw_forget_wisdom()
