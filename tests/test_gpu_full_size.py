"""Parity at the full sizes the benchmark is quoted on.

  * 1-D C2C complex128, N = 2^24 and N = 3 x 2^20, in place (four-step path);
  * batched C2C complex64, N = 1024 x 4096 transforms;
  * 3-D R2C float32 512^3 (real-packed row pass + two column passes).

Each is checked against the oracle where the oracle finishes in seconds,
against direct DFT sums X_k = sum_j x_j exp(-2 pi i jk/N) at sampled bins
(the definition, independent of the oracle; phases reduced exactly in integer
arithmetic before the exponential), and through size-independent properties:
Parseval's identity and the forward / backward round trip.

Tolerances (relative L2, or per bin relative to the spectrum's RMS): FP32
2e-7 * log2(N), FP64 1e-14 * log2(N); a sampled bin may be off by up to 8x
the transform-wide bound.
"""
import numpy as np
import pytest

import fftz
from gpu_util import rand_c, run_c2c, to_dev, tol
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def phase(n, idx, k):
    return np.exp(-2j * np.pi * ((idx * int(k)) % n) / n)


def direct_bins_1d(x, ks):
    n = x.size
    idx = np.arange(n, dtype=np.int64)
    return np.array([np.sum(x * phase(n, idx, k)) for k in ks])


@pytest.mark.parametrize("n", [1 << 24, 3 * (1 << 20)])
def test_large_fp64_in_place(torch_gpu, n):
    torch = torch_gpu
    print(f"full-size C2C fp64 n={n}", flush=True)
    rng = np.random.default_rng(24)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    d = to_dev(torch, x)
    f = fftz.Plan([(n, 1, 1)], [(1, n, n)], "d", 0, inplace=True, inp=d, out=d)
    b = fftz.Plan([(n, 1, 1)], [(1, n, n)], "d", 1, inplace=True, inp=d, out=d)
    assert "four-step" in f.info, f.info
    assert f.execute() == 0
    torch.cuda.synchronize()
    y = d.cpu().numpy()
    t = tol(n, "d")
    assert O.rel_l2(y, O.fft(x.reshape(1, n)).reshape(-1)) <= t
    ks = [0, 1, n // 2, n - 1] + [int(k) for k in rng.integers(0, n, 4)]
    rms = np.linalg.norm(y) / np.sqrt(n)
    assert np.max(np.abs(y[ks] - direct_bins_1d(x, ks))) <= 8 * t * rms
    energy = lambda v: (v.real ** 2 + v.imag ** 2).sum()  # pairwise summation
    assert abs(energy(y) / (n * energy(x)) - 1) <= t  # Parseval
    assert b.execute() == 0
    torch.cuda.synchronize()
    assert O.rel_l2(d.cpu().numpy() / n, x) <= 2 * t
    f.destroy()
    b.destroy()


def test_batched_n1024_fp32(torch_gpu):
    torch = torch_gpu
    n, b = 1024, 4096
    print(f"full-size batched C2C fp32 n={n} batch={b}", flush=True)
    x = rand_c((b, n), 7, "f")
    y, info = run_c2c(torch, x)
    ref = O.fft(x.astype(np.complex128)).reshape(b, n)
    t = tol(n, "f")
    assert O.rel_l2(y, ref) <= t, info
    per = np.linalg.norm(y - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert per.max() <= 4 * t  # every transform, not just the aggregate
    z, _ = run_c2c(torch, y, direction=1)
    assert O.rel_l2(z / n, x) <= 2 * t


def test_r2c_512_cubed_fp32(torch_gpu):
    torch = torch_gpu
    n, h = 512, 257
    print("full-size 3-D R2C fp32 512^3", flush=True)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((n, n, n), dtype=np.float32)
    dims = [(n, 1, 1), (n, n, h), (n, n * n, n * h)]
    vecs = [(1, n ** 3, n * n * h)]
    dx = to_dev(torch, x)
    dy = torch.empty((n, n, h), dtype=torch.complex64, device="cuda")
    p = fftz.Plan(dims, vecs, "f", 0, fft_type=1, inp=dx, out=dy)
    assert p.execute() == 0
    torch.cuda.synchronize()
    y = dy.cpu().numpy()
    p.destroy()
    del dx, dy
    # the same transform in float64 through the library
    dxd = to_dev(torch, x.astype(np.float64))
    dyd = torch.empty((n, n, h), dtype=torch.complex128, device="cuda")
    q = fftz.Plan(dims, vecs, "d", 0, fft_type=1, inp=dxd, out=dyd)
    assert q.execute() == 0
    torch.cuda.synchronize()
    yd = dyd.cpu().numpy()
    q.destroy()
    del dxd, dyd
    ntot = n ** 3
    tf, td = tol(ntot, "f"), tol(ntot, "d")
    assert O.rel_l2(y, yd) <= tf
    # direct sums at sampled bins (k0 along the innermost axis, k0 <= n/2)
    xd = x.reshape(-1, n).astype(np.float64)
    idx = np.arange(n, dtype=np.int64)
    rms = np.linalg.norm(yd) / np.sqrt(yd.size)
    for k0, k1, k2 in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (256, 256, 256), (3, 100, 200),
                       (129, 7, 511), (13, 400, 100)]:
        e0 = phase(n, idx, k0)
        v = (xd @ e0.real + 1j * (xd @ e0.imag)).reshape(n, n)  # [plane, row]
        s = (v @ phase(n, idx, k1)) @ phase(n, idx, k2)
        assert abs(yd[k2, k1, k0] - s) <= 8 * td * rms, (k0, k1, k2)
        assert abs(y[k2, k1, k0] - s) <= 8 * tf * rms, (k0, k1, k2)
    # Parseval over the half spectrum: bins 0 < k0 < n/2 stand for two bins
    e = (np.abs(yd[..., 0]) ** 2).sum() + (np.abs(yd[..., h - 1]) ** 2).sum() + \
        2 * (np.abs(yd[..., 1:h - 1]) ** 2).sum()
    assert abs(e / (ntot * np.sum(xd * xd)) - 1) <= td
