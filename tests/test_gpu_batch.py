"""GPU parity for uneven batches (many lengths in one persistent launch).

The reference plans and executes every (n, batch) item separately; here the
whole list runs through one work queue.  Checked against the oracle per item
(small lists: every item; the full 20,000-item workload: a seeded sample plus
every item through the round trip), bitwise run to run, and each item's
result alone equal to its result inside the mix."""
import numpy as np
import pytest

import fftz
import workloads
from gpu_util import tol
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def pack(items):
    off = 0
    out = []
    for it in items:
        out.append((it[0], it[1], off, off))
        off += it[0] * it[1]
    return out, off


def rand_dev(torch, total, seed, dt):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    re = torch.rand(total, dtype=torch.float64 if dt == torch.complex128 else torch.float32, device="cuda",
                    generator=g) * 2 - 1
    im = torch.rand(total, dtype=re.dtype, device="cuda", generator=g) * 2 - 1
    return torch.complex(re, im).to(dt)


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("direction", [0, 1])
def test_small_uneven_all_items(torch_gpu, prec, direction):
    torch = torch_gpu
    items, total = workloads.uneven_batch(300, seed=21, max_batch=8)
    if prec == "d":
        items, total = pack([it for it in items if it[0] <= 4096 or it[0] == 8192])
    dt = torch.complex64 if prec == "f" else torch.complex128
    x = rand_dev(torch, total, 5, dt)
    y = torch.empty_like(x)
    p = fftz.BatchPlan(items, prec, direction)
    assert p.execute_io(x, y) == 0
    torch.cuda.synchronize()
    xh, yh = x.cpu().numpy(), y.cpu().numpy()
    sign = -1 if direction == 0 else 1
    worst = 0.0
    for n, b, io, oo in items:
        ref = O.fft(xh[io:io + n * b].reshape(b, n).astype(np.complex128), sign)
        err = O.rel_l2(yh[oo:oo + n * b], ref.reshape(-1))
        assert err <= tol(n, prec), (n, b, err)
        worst = max(worst, err / tol(n, prec))
    print(f"worst err/tol {worst:.3f}; {p.info}")
    p.destroy()


EDGE_ITEMS = [(1, 5), (2, 3), (3, 1), (5, 4), (7, 2), (8, 6), (13, 3), (16, 2), (17, 4), (31, 1), (97, 3),
              (1, 1), (101, 2), (1024, 2), (4099, 1), (2, 64), (15, 9)]


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("direction", [0, 1])
def test_uneven_edge_lengths(torch_gpu, prec, direction):
    """Length-1 items (copied), one-stage lengths (2..16, register-only),
    small primes whose Bluestein length is the smallest two-stage one
    (17 -> 33), and long ones, mixed in one queue; every item against the
    oracle."""
    torch = torch_gpu
    items, total = pack(EDGE_ITEMS)
    dt = torch.complex64 if prec == "f" else torch.complex128
    x = rand_dev(torch, total, 9, dt)
    y = torch.full_like(x, float("nan"))
    p = fftz.BatchPlan(items, prec, direction)
    assert p.execute_io(x, y) == 0
    torch.cuda.synchronize()
    xh, yh = x.cpu().numpy(), y.cpu().numpy()
    sign = -1 if direction == 0 else 1
    for n, b, io, oo in items:
        ref = O.fft(xh[io:io + n * b].reshape(b, n).astype(np.complex128), sign)
        err = O.rel_l2(yh[oo:oo + n * b], ref.reshape(-1))
        assert err <= tol(max(n, 2), prec), (n, b, err, p.info)
    p.destroy()


def test_uneven_full_size(torch_gpu):
    """20,000 items, lengths from seed 3 (70% 7-smooth in [16, 16384], 30%
    primes in [101, 8191]), batch 1..64, complex64, packed contiguously."""
    torch = torch_gpu
    items, total = workloads.uneven_batch(20000, seed=3)
    x = rand_dev(torch, total, 3, torch.complex64)
    y = torch.empty_like(x)
    fwd = fftz.BatchPlan(items, "f", 0)
    print(fwd.info)
    assert fwd.execute_io(x, y) == 0
    torch.cuda.synchronize()
    # sampled items against the oracle
    rng = np.random.default_rng(99)
    primes = set(workloads.primes())
    pick = list(rng.choice(len(items), 48, replace=False))
    pick += [i for i, it in enumerate(items) if it[0] in primes][:8]
    pick += [int(np.argmax([it[0] for it in items]))]
    for i in pick:
        n, b, io, oo = items[i]
        xs = x[io:io + n * b].cpu().numpy().astype(np.complex128).reshape(b, n)
        ys = y[oo:oo + n * b].cpu().numpy()
        err = O.rel_l2(ys, O.fft(xs).reshape(-1))
        assert err <= tol(n, "f"), (n, b, err)
    # bitwise reproducible
    y2 = torch.empty_like(x)
    assert fwd.execute_io(x, y2) == 0
    torch.cuda.synchronize()
    assert torch.equal(y.view(torch.float32), y2.view(torch.float32))
    del y2
    # each item alone equals its result in the mix
    for i in pick[:6]:
        n, b, io, oo = items[i]
        xa = x[io:io + n * b].clone()
        ya = torch.empty_like(xa)
        pa = fftz.BatchPlan([(n, b, 0, 0)], "f", 0)
        assert pa.execute_io(xa, ya) == 0
        torch.cuda.synchronize()
        assert torch.equal(ya.view(torch.float32), y[oo:oo + n * b].view(torch.float32))
        pa.destroy()
    # every item through the round trip: backward(forward(x)) / n == x
    bwd = fftz.BatchPlan(items, "f", 1)
    back = torch.empty_like(x)
    assert bwd.execute_io(y, back) == 0
    torch.cuda.synchronize()
    ns = torch.tensor([it[0] for it in items], dtype=torch.float32, device="cuda")
    reps = torch.tensor([it[0] * it[1] for it in items], dtype=torch.int64, device="cuda")
    scale = torch.repeat_interleave(1.0 / ns, reps)
    back.mul_(scale)
    del scale
    back.sub_(x)
    rt = (torch.linalg.vector_norm(back) / torch.linalg.vector_norm(x)).item()
    print(f"round trip rel-L2 over all {len(items)} items: {rt:.3e}")
    assert rt <= 2 * 2e-7 * 14
    fwd.destroy()
    bwd.destroy()


def test_uneven_autotune(torch_gpu):
    torch = torch_gpu
    items, total = workloads.uneven_batch(300, seed=21, max_batch=8)
    x = rand_dev(torch, total, 5, torch.complex64)
    y = torch.empty_like(x)
    p = fftz.BatchPlan(items, "f", 0)
    assert p.autotune(x.clone(), torch.empty_like(x), reps=2) == 0
    assert p.execute_io(x, y) == 0
    torch.cuda.synchronize()
    xh, yh = x.cpu().numpy(), y.cpu().numpy()
    for n, b, io, oo in items[:40]:
        ref = O.fft(xh[io:io + n * b].reshape(b, n).astype(np.complex128))
        assert O.rel_l2(yh[oo:oo + n * b], ref.reshape(-1)) <= tol(n, "f"), (n, b)
    print(p.info)
    p.destroy()
