"""Transforms whose length exceeds one LDS tile, beyond the plain 1-D
four-step: multi-pass Bluestein (prime lengths whose extended length M is
above the single-pass limit: 16384 complex64, 8192 complex128), long axes of
N-D problems, and long 1-D transforms with several batch axes.

Reference behaviour being matched: Bluestein for any n
(selector/complex/selector_bluestein_dft.c:20-105, execute
core/solvers/complex/bluestein_solver_dft.c:137-234); every axis of an N-D
problem as a full 1-D sub-plan (core/solvers/complex/ndim_solver_dft.c:17-121);
batches over any vec rank (core/solvers/complex/batched_solver_dft.c:33-110).

Each case is compared with the oracle on the same seeded input (relative L2,
FP32 2e-7 log2 N, FP64 1e-14 log2 N) and checked through the forward /
backward round trip.
"""
import numpy as np
import pytest

import fftz
from gpu_util import rand_c, run_c2c, run_desc, to_dev, tol
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def check_roundtrip(torch, x, y, prec):
    n = x.shape[-1]
    z, _ = run_c2c(torch, y, direction=1)
    assert O.rel_l2(z / n, x) <= 2 * tol(n, prec)


@pytest.mark.parametrize("n,prec,b", [(16411, "f", 2), (65537, "d", 1), (131071, "f", 1), (8209, "d", 3)])
def test_bluestein_multi_pass(torch_gpu, n, prec, b):
    print(f"multi-pass Bluestein n={n} prec={prec} batch={b}", flush=True)
    x = rand_c((b, n), n, prec)
    y, info = run_c2c(torch_gpu, x)
    assert "multi-pass" in info, info
    ref = O.fft(x.astype(np.complex128))
    assert O.rel_l2(y, ref) <= tol(n, prec), info
    yb, _ = run_c2c(torch_gpu, x, direction=1)
    assert O.rel_l2(yb, O.fft(x.astype(np.complex128), +1)) <= tol(n, prec)
    check_roundtrip(torch_gpu, x, y, prec)


def test_bluestein_multi_pass_strided_inplace(torch_gpu):
    # two interleaved sequences (element stride 2, batch stride 1), in place
    n, prec = 16411, "f"
    print(f"multi-pass Bluestein strided in place n={n}", flush=True)
    x = rand_c(2 * n, 3, prec)
    out = run_desc(torch_gpu, x, 2 * n, [(n, 2, 2)], [(2, 1, 1)], prec, 0, inplace=True)
    ref = x.astype(np.complex128).copy()
    O.c2c_strided(x.astype(np.complex128), ref, [(n, 2, 2)], [(2, 1, 1)], -1)
    assert O.rel_l2(out, ref) <= tol(n, prec)


def test_prime_factor_2p20_times_17(torch_gpu):
    """complex64 N = 2^20 * 17: a prime factor above 13 at a length whose
    extended length (35,654,850) needs the multi-pass form."""
    torch = torch_gpu
    n = (1 << 20) * 17
    print(f"complex64 n=2^20*17 ({n})", flush=True)
    rng = np.random.default_rng(17)
    x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)
    d = to_dev(torch, x)
    o = torch.empty_like(d)
    p = fftz.Plan([(n, 1, 1)], [(1, n, n)], "f", 0, inp=d, out=o)
    assert "multi-pass" in p.info, p.info
    assert p.execute() == 0
    torch.cuda.synchronize()
    y = o.cpu().numpy()
    p.destroy()
    ref = O.fft(x.astype(np.complex128).reshape(1, n)).reshape(-1)
    assert O.rel_l2(y, ref) <= tol(n, "f")
    q = fftz.Plan([(n, 1, 1)], [(1, n, n)], "f", 1, inp=o, out=d)
    assert q.execute() == 0
    torch.cuda.synchronize()
    q.destroy()
    assert O.rel_l2(d.cpu().numpy() / n, x) <= 2 * tol(n, "f")


LONG = [
    ([(8, 1, 1), (32768, 8, 8)], [(1, 1, 1)], "f"),                     # 2-D 8 x 32768, long outer axis
    ([(32768, 1, 1), (8, 32768, 32768)], [(1, 1, 1)], "f"),             # long inner axis
    ([(65536, 1, 1), (4, 65536, 65536), (4, 4 * 65536, 4 * 65536)], [(1, 1, 1)], "f"),  # 3-D 4 x 4 x 65536
    ([(20000, 1, 1), (3, 20000, 20000)], [(2, 60000, 60000)], "d"),     # FP64 long axis, batched
    ([(1 << 20, 1, 1)], [(3, 1 << 20, 1 << 20), (2, 3 << 20, 3 << 20)], "f"),  # 2^20, two batch axes
    ([(1 << 14, 1, 1)], [(3, 1 << 14, 1 << 14), (2, 3 << 14, 3 << 14), (2, 6 << 14, 6 << 14)], "d"),
]


@pytest.mark.parametrize("dims,vecs,prec", LONG)
@pytest.mark.parametrize("direction", [0, 1])
def test_long_axes(torch_gpu, dims, vecs, prec, direction):
    print(f"long axis dims={dims} vecs={vecs} prec={prec} dir={direction}", flush=True)
    ntot = 1
    for d in dims:
        ntot *= d[0]
    for v in vecs:
        ntot *= v[0]
    x = rand_c(ntot, ntot % 1000, prec)
    out = run_desc(torch_gpu, x, ntot, dims, vecs, prec, direction)
    ref = np.zeros(ntot, np.complex128)
    O.c2c_strided(x.astype(np.complex128), ref, dims, vecs, -1 if direction == 0 else 1)
    n = 1
    for d in dims:
        n *= d[0]
    assert O.rel_l2(out, ref) <= tol(n, prec)
