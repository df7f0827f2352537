"""GPU tests of the allocation-free execute (aoclfftz_gpu_execute_noalloc).

The entry point runs the same kernels as aoclfftz_execute_io, with the work
memory of four-step / multi-pass plans and the uneven batch's queue counters
taken from a static per-GPU scratch array (64 MiB of work memory) instead of
an allocation.  Each case compares its output bit for bit with
aoclfftz_execute_io on the same input, over two calls in a row where the
second call's input is smaller (the array still holds the first call's
contents), plus the oracle on one case.  Plans needing more than the array
holds, and host buffers, are refused without launching anything.
"""
import numpy as np
import pytest

import fftz
import workloads
from gpu_util import rand_c, to_dev, tol
from oracle import oracle as O

pytestmark = pytest.mark.gpu

INVALID_INPUT = -3


def same_bits(a, b):
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


def run_pair(torch, dims, vecs, x, prec, direction=0, inplace=False):
    """(noalloc output, execute_io output) of one plan on one input."""
    dx = to_dev(torch, x)
    p = fftz.Plan(dims, vecs, prec, direction, inplace=inplace)
    if inplace:
        a = dx.clone()
        b = dx.clone()
        assert p.execute_noalloc(a, a) == 0, p.info
        assert p.execute_io(b, b) == 0
    else:
        a = torch.full_like(dx, float("nan"))
        b = torch.empty_like(dx)
        assert p.execute_noalloc(dx, a) == 0, p.info
        assert p.execute_io(dx, b) == 0
    torch.cuda.synchronize()
    info = p.info
    p.destroy()
    return a.cpu().numpy(), b.cpu().numpy(), info


def test_noalloc_four_step_two_calls(torch_gpu):
    torch = torch_gpu
    for n, b in ((1 << 21, 2), (1 << 18, 1)):  # second call smaller
        print(f"noalloc four-step complex64 n={n} batch={b}", flush=True)
        x = rand_c((b, n), n + 1, "f")
        got, want, info = run_pair(torch, [(n, 1, 1)], [(b, n, n)], x, "f")
        assert "four-step" in info, info
        assert same_bits(got, want), info
    ref = O.fft(x.astype(np.complex128))
    assert O.rel_l2(got, ref) <= tol(n, "f")


def test_noalloc_bluestein_multi_pass_two_calls(torch_gpu):
    torch = torch_gpu
    for n in (131071, 16411):
        print(f"noalloc multi-pass Bluestein complex64 n={n} backward", flush=True)
        x = rand_c((1, n), n, "f")
        got, want, info = run_pair(torch, [(n, 1, 1)], [(1, n, n)], x, "f", direction=1)
        assert "multi-pass" in info, info
        assert same_bits(got, want), info


def test_noalloc_largest_work_in_place_fp64(torch_gpu):
    # 2^22 complex128 in place: a 64 MiB work buffer, the array's whole size
    torch = torch_gpu
    n = 1 << 22
    print(f"noalloc four-step complex128 n={n} in place", flush=True)
    x = rand_c((1, n), 7, "d")
    got, want, info = run_pair(torch, [(n, 1, 1)], [(1, n, n)], x, "d", inplace=True)
    assert same_bits(got, want), info


def test_noalloc_uneven_batch_two_calls(torch_gpu):
    torch = torch_gpu
    for count, seed in ((120, 31), (20, 32)):
        print(f"noalloc uneven batch: {count} items", flush=True)
        items, total = workloads.uneven_batch(count, seed=seed, max_batch=4)
        x = to_dev(torch, rand_c(total, seed, "f"))
        p = fftz.BatchPlan(items, "f", 0)
        a = torch.full_like(x, float("nan"))
        b = torch.empty_like(x)
        assert p.execute_noalloc(x, a) == 0
        assert p.execute_io(x, b) == 0
        torch.cuda.synchronize()
        assert same_bits(a.cpu().numpy(), b.cpu().numpy()), p.info
        p.destroy()


def test_noalloc_stream_mode(torch_gpu):
    torch = torch_gpu
    n, b = 1 << 19, 2
    print(f"noalloc on a bound stream complex64 n={n}", flush=True)
    x = to_dev(torch, rand_c((b, n), 11, "f"))
    p = fftz.Plan([(n, 1, 1)], [(b, n, n)], "f", 0)
    s = torch.cuda.Stream()
    assert p.set_stream(s) == 0
    a = torch.empty_like(x)
    bb = torch.empty_like(x)
    torch.cuda.synchronize()
    assert p.execute_noalloc(x, a) == 0
    assert p.execute_io(x, bb) == 0
    s.synchronize()
    assert same_bits(a.cpu().numpy(), bb.cpu().numpy())
    p.destroy()


def test_noalloc_refuses_oversize_and_host_buffers(torch_gpu):
    torch = torch_gpu
    # 2^23 complex128 in place needs a 128 MiB work buffer: refused, nothing runs
    n = 1 << 23
    print(f"noalloc refuses complex128 n={n}", flush=True)
    x = torch.zeros(n, dtype=torch.complex128, device="cuda")
    p = fftz.Plan([(n, 1, 1)], [(1, n, n)], "d", 0, inplace=True)
    assert p.execute_noalloc(x, x) == INVALID_INPUT
    torch.cuda.synchronize()
    assert torch.count_nonzero(x).item() == 0
    p.destroy()
    # host (numpy) buffers would need staging memory: refused
    h = rand_c((1, 1024), 1, "f")
    ho = np.zeros_like(h)
    q = fftz.Plan([(1024, 1, 1)], [(1, 1024, 1024)], "f", 0)
    assert q.execute_noalloc(h, ho) == INVALID_INPUT
    assert not ho.any()
    q.destroy()
