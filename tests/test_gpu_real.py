"""GPU parity for real transforms (real-packed R2C / C2R) through the aoclfftz
C ABI: 1-D, 2-D, 3-D, batched, in-place and out-of-place, FP32 and FP64,
against the FP64 oracle (oracle.r2c restates the definition; oracle.r2c_packed
restates the reference's recombine, core/kernels/non_dft/real_packed/
packed_rdft_common.h:19-60, and both are pinned against numpy in
tests/test_oracle.py)."""
import numpy as np
import pytest

import fftz
from gpu_util import to_dev, tol
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def r2c_dims(shape, inplace):
    n0 = shape[-1]
    h = n0 // 2 + 1
    dims = [(n0, 1, 1)]
    si = 2 * h if inplace else n0
    so = h
    for n in reversed(shape[:-1]):
        dims.append((n, si, so))
        si *= n
        so *= n
    return dims, si, so


# odd innermost lengths and even ones whose half has no radix plan take the
# full-length path (real promoted to complex; Hermitian extension for C2R),
# Bluestein for prime factors above 13
FULL_LENGTH = [(1,), (7,), (9,), (101,), (202,), (2187,), (9, 15), (5, 7, 11), (4, 101)]


@pytest.mark.parametrize("shape", [(2,), (16,), (1000,), (4096,), (16384,), (8, 16), (64, 48), (12, 10, 20),
                                   (32, 32, 32)] + FULL_LENGTH)
@pytest.mark.parametrize("prec", ["f", "d"])
def test_r2c_c2r_out_of_place(torch_gpu, shape, prec):
    torch = torch_gpu
    if prec == "d" and shape == (16384,):
        shape = (8192,)
    dims, si, so = r2c_dims(shape, False)
    b = 3
    rt = np.float32 if prec == "f" else np.float64
    ct = np.complex64 if prec == "f" else np.complex128
    rng = np.random.default_rng(len(shape) * 100 + shape[-1])
    x = rng.standard_normal((b,) + shape).astype(rt)
    dx = to_dev(torch, x.reshape(-1))
    dy = torch.zeros(b * so, dtype=getattr(torch, np.dtype(ct).name), device="cuda")
    p = fftz.Plan(dims, [(b, si, so)], prec, 0, fft_type=1, inp=dx, out=dy)
    assert p.execute() == 0
    torch.cuda.synchronize()
    y = dy.cpu().numpy()
    ref = O.rfftn(x.astype(np.float64), batch_dims=1)
    n = int(np.prod(shape))
    assert O.rel_l2(y, ref.reshape(-1)) <= tol(n, prec)
    p.destroy()
    # C2R back (input preserved)
    dback = torch.zeros(b * si, dtype=getattr(torch, np.dtype(rt).name), device="cuda")
    dy_copy = dy.clone()
    cdims = [(dims[0][0], dims[0][2], dims[0][1])] + [(d[0], d[2], d[1]) for d in dims[1:]]
    q = fftz.Plan(cdims, [(b, so, si)], prec, 1, fft_type=1, inp=dy, out=dback)
    assert q.execute() == 0
    torch.cuda.synchronize()
    back = dback.cpu().numpy().reshape((b,) + shape)
    assert O.rel_l2(back / n, x.astype(np.float64)) <= 2 * tol(n, prec)
    assert torch.equal(dy, dy_copy)
    q.destroy()


@pytest.mark.parametrize("shape", [(64,), (8, 30), (6, 10, 16), (9,), (6, 10, 15), (3, 101)])
def test_r2c_inplace(torch_gpu, shape):
    torch = torch_gpu
    dims, si, so = r2c_dims(shape, True)
    b = 2
    rng = np.random.default_rng(4)
    n0 = shape[-1]
    h = n0 // 2 + 1
    x = rng.standard_normal((b,) + shape)
    padded = np.zeros((b,) + shape[:-1] + (2 * h,))
    padded[..., :n0] = x
    buf = to_dev(torch, padded.reshape(-1))
    p = fftz.Plan(dims, [(b, si, so)], "d", 0, fft_type=1, inplace=True, inp=buf, out=buf)
    assert p.execute() == 0
    torch.cuda.synchronize()
    y = buf.cpu().numpy().view(np.complex128)
    ref = O.rfftn(x, batch_dims=1)
    assert O.rel_l2(y, ref.reshape(-1)) <= tol(int(np.prod(shape)), "d")
    p.destroy()


def test_cfg4a_shape_r2c_512_cubed_properties(torch_gpu):
    """512^3 float32 R2C at full size: sum of 8 random-frequency sinusoids
    plus 1e-3 noise.  Checked through size-independent properties: the
    sinusoid bins carry the analytic amplitude, the C2R round trip returns
    the input, and a sample of bins matches a direct sum on the CPU."""
    torch = torch_gpu
    n = 512
    rng = np.random.default_rng(2024)
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    freqs = rng.integers(0, n, size=(8, 3))
    amps = rng.uniform(0.5, 2.0, 8)
    i = torch.arange(n, device="cuda", dtype=torch.float64)
    x = torch.zeros((n, n, n), dtype=torch.float64, device="cuda")
    for (f0, f1, f2), a in zip(freqs, amps):
        ph = (2 * np.pi / n) * (f0 * i[:, None, None] + f1 * i[None, :, None] + f2 * i[None, None, :])
        x += a * torch.cos(ph)
    x += 1e-3 * torch.randn((n, n, n), dtype=torch.float64, device="cuda", generator=g)
    xf = x.to(torch.float32).contiguous()
    del x
    h = n // 2 + 1
    dims = [(n, 1, 1), (n, n, h), (n, n * n, n * h)]
    y = torch.empty((n, n, h), dtype=torch.complex64, device="cuda")
    p = fftz.Plan(dims, [(1, n ** 3, n * n * h)], "f", 0, fft_type=1, inp=xf, out=y)
    assert p.execute() == 0
    torch.cuda.synchronize()
    # sampled bins against a direct FP64 sum over the FP32 input
    xs = xf.double()
    idx = [(int(a), int(b), int(c)) for a, b, c in rng.integers(0, [n, n, h], size=(6, 3))]
    for (f0, f1, f2), a in zip(freqs[:3], amps[:3]):
        k2 = f2 if f2 <= n // 2 else (n - f2)
        k0, k1 = (f0, f1) if f2 <= n // 2 else ((n - f0) % n, (n - f1) % n)
        idx.append((int(k0), int(k1), int(k2)))
    e = torch.arange(n, device="cuda", dtype=torch.float64)
    xc = xs.to(torch.complex128)
    # per-bin FP32 error: accumulated rounding at the rms level |X|_rms =
    # ||x||_2 (Parseval) plus the bin's own representation error (peaks reach
    # ~4e7, where one float32 ulp is 4)
    xnorm = torch.linalg.vector_norm(xs).item()
    for (k0, k1, k2) in idx:
        w0 = torch.exp(-2j * np.pi * k0 * e / n)
        w1 = torch.exp(-2j * np.pi * k1 * e / n)
        w2 = torch.exp(-2j * np.pi * k2 * e / n)
        direct = torch.matmul(torch.matmul(torch.matmul(xc, w2), w1), w0).item()
        got = complex(y[k0, k1, k2].item())
        bound = 2e-7 * 27 * xnorm * 3 + 2.4e-7 * abs(direct)
        assert abs(got - direct) <= bound, ((k0, k1, k2), got, direct, bound)
    del xc
    # C2R round trip
    back = torch.empty((n, n, n), dtype=torch.float32, device="cuda")
    cdims = [(n, 1, 1), (n, h, n), (n, n * h, n * n)]
    q = fftz.Plan(cdims, [(1, n * n * h, n ** 3)], "f", 1, fft_type=1, inp=y, out=back)
    assert q.execute() == 0
    torch.cuda.synchronize()
    err = (torch.linalg.vector_norm(back.double() / n ** 3 - xs) / torch.linalg.vector_norm(xs)).item()
    assert err < 2e-7 * 27 * 2
    p.destroy()
    q.destroy()
