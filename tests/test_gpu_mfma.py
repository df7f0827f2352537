"""GPU parity for the matrix-core radix-16 pass (aoclfftz_gpu_set_compute_mode)
on complex64 lengths 256 and 4096, against the FP64 oracle.

Relative-L2 bounds, stated per operand precision:
  FP32 operands (mode 1) and scaled fp16 hi+lo (mode 3): 2e-7 * log2 N, the
  plain FP32 bound (fp16 pairs carry 22 significant bits; each line is scaled
  by a power of two into the fp16 range);
  bf16 hi+lo (mode 2): 1e-5 * log2 N (16 significant bits per operand, the
  lo*lo product dropped).
"""
import numpy as np
import pytest

import fftz
import workloads
from gpu_util import rand_c, to_dev, tol
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MODES = {1: "mfma-f32", 2: "mfma-bf16x3", 3: "mfma-f16x2"}


def mode_tol(n, mode):
    return 1e-5 * np.log2(n) if mode == 2 else tol(n, "f")


def run(torch, x, mode, direction=0):
    b, n = x.shape
    dx = to_dev(torch, x)
    dy = torch.empty_like(dx)
    p = fftz.Plan([(n, 1, 1)], [(b, n, n)], "f", direction, inp=dx, out=dy)
    assert p.set_compute_mode(mode) == 0
    info = p.info
    assert p.execute() == 0
    torch.cuda.synchronize()
    y = dy.cpu().numpy()
    p.destroy()
    return y, info


@pytest.mark.parametrize("mode", [1, 2, 3])
@pytest.mark.parametrize("n,b", [(256, 37), (4096, 11)])
@pytest.mark.parametrize("direction", [0, 1])
def test_mfma_uniform(torch_gpu, mode, n, b, direction):
    x = rand_c((b, n), n + mode, "f")
    y, info = run(torch_gpu, x, mode, direction)
    assert MODES[mode] in info, info
    ref = O.fft(x.astype(np.complex128), -1 if direction == 0 else 1)
    err = O.rel_l2(y, ref)
    print(f"mode {mode} n={n} dir={direction} rel-L2 {err:.3e} (bound {mode_tol(n, mode):.1e})")
    assert err <= mode_tol(n, mode)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_mfma_sinusoids_vs_plain(torch_gpu, mode):
    """Lines of 8 tones + 1e-3 noise (the precision workload's input): every
    variant against the FP64 oracle."""
    x = workloads.sinusoids(4096, 64, seed=4)
    y, info = run(torch_gpu, x, mode)
    err = O.rel_l2(y, O.fft(x.astype(np.complex128)))
    print(f"sinusoids mode {mode}: rel-L2 {err:.3e}  {info}")
    assert err <= (mode_tol(4096, mode) if mode else tol(4096, "f"))


def test_mfma_bitwise_and_strided(torch_gpu):
    torch = torch_gpu
    # lines 4100 elements apart inside a padded buffer, output contiguous
    n, b, dist = 4096, 9, 4100
    x = rand_c((b * dist,), 5, "f")
    dx = to_dev(torch, x)
    outs = []
    for _ in range(2):
        dy = torch.zeros(b * n, dtype=torch.complex64, device="cuda")
        p = fftz.Plan([(n, 1, 1)], [(b, dist, n)], "f", 0, inp=dx, out=dy)
        assert p.set_compute_mode(3) == 0
        assert p.execute() == 0
        torch.cuda.synchronize()
        outs.append(dy.cpu().numpy())
        p.destroy()
    assert np.array_equal(outs[0].view(np.float32), outs[1].view(np.float32))
    lines = np.stack([x[i * dist:i * dist + n] for i in range(b)]).astype(np.complex128)
    assert O.rel_l2(outs[0], O.fft(lines).reshape(-1)) <= tol(n, "f")


def test_mfma_mode_rejected_where_not_applicable(torch_gpu):
    torch = torch_gpu
    x = to_dev(torch, rand_c((4, 1000), 1, "f"))
    y = torch.empty_like(x)
    p = fftz.Plan([(1000, 1, 1)], [(4, 1000, 1000)], "f", 0, inp=x, out=y)
    assert p.set_compute_mode(2) == -3  # AOCLFFTZ_INVALID_INPUT, plan unchanged
    assert p.set_compute_mode(7) == -3
    assert p.execute() == 0
    torch.cuda.synchronize()
    p.destroy()
    xd = to_dev(torch, rand_c((4, 4096), 1, "d"))
    q = fftz.Plan([(4096, 1, 1)], [(4, 4096, 4096)], "d", 0, inp=xd, out=torch.empty_like(xd))
    assert q.set_compute_mode(1) == -3  # complex128 plans have no matrix-core pass
    q.destroy()


def test_mfma_full_size_4096x16384(torch_gpu):
    """N = 4096, batch 16384 at full size: sampled lines against the oracle
    and the forward/backward round trip of every line."""
    torch = torch_gpu
    n, b = 4096, 16384
    g = torch.Generator(device="cuda")
    g.manual_seed(8)
    x = torch.empty(b * n, dtype=torch.complex64, device="cuda")
    x.view(torch.float32).uniform_(-1.0, 1.0, generator=g)
    y = torch.empty_like(x)
    back = torch.empty_like(x)
    for mode in (1, 2, 3):
        f = fftz.Plan([(n, 1, 1)], [(b, n, n)], "f", 0, inp=x, out=y)
        r = fftz.Plan([(n, 1, 1)], [(b, n, n)], "f", 1, inp=y, out=back)
        assert f.set_compute_mode(mode) == 0 and r.set_compute_mode(mode) == 0
        assert f.execute() == 0 and r.execute() == 0
        torch.cuda.synchronize()
        rows = [0, 1, 4095, 8191, b - 1]
        xs = x.view(b, n)[rows].cpu().numpy().astype(np.complex128)
        err = O.rel_l2(y.view(b, n)[rows].cpu().numpy(), O.fft(xs))
        rt = (torch.linalg.vector_norm(back / n - x) / torch.linalg.vector_norm(x)).item()
        print(f"mode {mode}: sampled rel-L2 {err:.3e}, round trip {rt:.3e}")
        assert err <= mode_tol(n, mode)
        assert rt <= 2 * mode_tol(n, mode)
        f.destroy()
        r.destroy()


@pytest.mark.parametrize("n,b", [(256, 37), (4096, 5)])
def test_mfma_fp16_lines_of_very_different_magnitude(torch_gpu, n, b):
    """Scaled fp16 operands are scaled per line: lines whose magnitudes span
    1e-12 .. 1e+12 each meet the FP32 bound on their own (a per-tile or global
    scale would leave the small lines to fp16 subnormals)."""
    x = rand_c((b, n), 77, "f")
    mags = np.logspace(-12, 12, b).astype(np.float32)
    x = (x * mags[:, None]).astype(np.complex64)
    y, info = run(torch_gpu, x, 3)
    assert MODES[3] in info, info
    ref = O.fft(x.astype(np.complex128))
    worst = max(O.rel_l2(y[i], ref[i]) for i in range(b))
    print(f"n={n}: worst per-line rel-L2 {worst:.3e}")
    assert worst <= tol(n, "f")
