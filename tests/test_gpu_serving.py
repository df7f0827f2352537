"""GPU tests of the serving runtime (aoclfftz_gpu_server_*).

A mixed stream of requests -- 1-D C2C of smooth and prime lengths, 2-D R2C,
3-D C2C, some complex128 -- goes through one server with a small HBM budget
and few slots, so requests queue, are admitted as others finish and overlap
on the device.  Every result must equal, bit for bit, the same problem run
alone through the ordinary setup / execute path; two results are also checked
against the oracle.  Admission must never exceed the budget.
"""
import numpy as np
import pytest

import fftz
from gpu_util import tol
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def request_mix():
    """(dims, vecs, precision, fft_type, in_elems, out_elems, in_dtype, out_dtype)"""
    reqs = []

    def c2c1(n, b, prec):
        ct = "complex64" if prec == "f" else "complex128"
        reqs.append(([(n, 1, 1)], [(b, n, n)], prec, 0, n * b, n * b, ct, ct))

    def r2c2(n, b):
        h = n // 2 + 1
        reqs.append(([(n, 1, 1), (n, n, h)], [(b, n * n, n * h)], "f", 1, n * n * b, n * h * b, "float32",
                     "complex64"))

    def c2c3(n, b, prec):
        ct = "complex64" if prec == "f" else "complex128"
        reqs.append(([(n, 1, 1), (n, n, n), (n, n * n, n * n)], [(b, n ** 3, n ** 3)], prec, 0, n ** 3 * b,
                     n ** 3 * b, ct, ct))

    for k in range(3):
        c2c1(64, 8 + k, "f")
        c2c1(1000, 3, "f")
        c2c1(4096, 2 + k, "f")
        c2c1(101, 4, "f")           # Bluestein
        c2c1(1 << 16, 2, "f")
        r2c2(128, 2)
        c2c3(32, 1, "f")
    c2c1(1 << 20, 1, "f")           # four-step
    r2c2(512, 1)
    c2c3(64, 1, "f")
    c2c1(4096, 2, "d")
    c2c3(32, 1, "d")
    c2c1(1 << 20, 1, "d")
    return reqs


def make_inputs(torch, reqs, seed=3):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    xs = []
    for dims, vecs, prec, ftype, nin, nout, idt, odt in reqs:
        if idt.startswith("complex"):
            rt = torch.float32 if idt == "complex64" else torch.float64
            x = torch.complex(torch.rand(nin, dtype=rt, device="cuda", generator=g) * 2 - 1,
                              torch.rand(nin, dtype=rt, device="cuda", generator=g) * 2 - 1)
        else:
            x = torch.rand(nin, dtype=getattr(torch, idt), device="cuda", generator=g) * 2 - 1
        xs.append(x)
    return xs


def test_serving_mix_equals_alone(torch_gpu):
    torch = torch_gpu
    reqs = request_mix()
    xs = make_inputs(torch, reqs)
    # alone: the ordinary setup / execute path, one at a time
    alone = []
    for (dims, vecs, prec, ftype, nin, nout, idt, odt), x in zip(reqs, xs):
        y = torch.empty(nout, dtype=getattr(torch, odt), device="cuda")
        p = fftz.Plan(dims, vecs, prec, 0, fft_type=ftype, inp=x, out=y)
        assert p.execute() == 0
        torch.cuda.synchronize()
        alone.append(y)
        p.destroy()
    # in the mix: 64 MiB budget (the largest request, 2^20 complex128 with its
    # four-step work buffer, needs 48 MiB), 4 slots -> queueing and overlap
    budget = 64 << 20
    srv = fftz.Server(budget, max_inflight=4)
    outs = [torch.full((r[5],), 7.0, dtype=getattr(torch, r[7]), device="cuda") for r in reqs]
    ids = []
    for (dims, vecs, prec, ftype, nin, nout, idt, odt), x, y in zip(reqs, xs, outs):
        rid = srv.submit(dims, vecs, prec, 0, fft_type=ftype, inp=x, out=y)
        assert rid >= 0
        ids.append(rid)
    lat = []
    for rid in ids:
        rc, t = srv.wait(rid)
        assert rc == 0
        lat.append(t)
    st = srv.stats()
    print(f"{len(reqs)} requests: {st}")
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert 0 < st["peak_footprint_bytes"] <= budget
    assert st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["plans_resident"] < len(reqs)  # repeated problems share one plan
    for i, (a, b) in enumerate(zip(alone, outs)):
        assert torch.equal(a.view(torch.uint8), b.view(torch.uint8)), (i, reqs[i][:4])
    # two results against the oracle
    for i in (1, 3):
        dims, vecs, prec, ftype, nin, nout, idt, odt = reqs[i]
        n, b = dims[0][0], vecs[0][0]
        xh = xs[i].cpu().numpy().astype(np.complex128).reshape(b, n)
        assert O.rel_l2(outs[i].cpu().numpy(), O.fft(xh).reshape(-1)) <= tol(n, prec)
    srv.destroy()


def test_serving_rejects_oversize_and_invalid(torch_gpu):
    torch = torch_gpu
    srv = fftz.Server(1 << 20, max_inflight=2)
    x = torch.zeros(1 << 18, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    # 2 x 2 MiB extents > 1 MiB budget
    assert srv.submit([(1 << 18, 1, 1)], [(1, 1 << 18, 1 << 18)], "f", 0, inp=x, out=y) == -2
    # zero length: invalid descriptor
    assert srv.submit([(0, 1, 1)], [(1, 1, 1)], "f", 0, inp=x, out=y) == -1
    small = x[:1024]
    rid = srv.submit([(1024, 1, 1)], [(1, 1024, 1024)], "f", 0, inp=small, out=y[:1024])
    assert rid >= 0
    assert srv.wait(rid)[0] == 0
    assert srv.test(rid) == -1  # forgotten after wait
    srv.destroy()


def test_serving_concurrent_submitters(torch_gpu):
    """Four host threads submit and wait concurrently (the C calls release the
    GIL); every result equals the same problem run alone."""
    import threading
    torch = torch_gpu
    n, b = 4096, 4
    g = torch.Generator(device="cuda")
    g.manual_seed(12)
    xs = [torch.complex(torch.rand(n * b, device="cuda", generator=g), torch.rand(n * b, device="cuda", generator=g))
          for _ in range(32)]
    ref = []
    for x in xs:
        y = torch.empty_like(x)
        p = fftz.Plan([(n, 1, 1)], [(b, n, n)], "f", 0, inp=x, out=y)
        assert p.execute() == 0
        torch.cuda.synchronize()
        ref.append(y)
        p.destroy()
    outs = [torch.empty_like(x) for x in xs]
    srv = fftz.Server(256 << 20, max_inflight=6)
    errors = []

    def worker(k):
        try:
            for i in range(k, len(xs), 4):
                rid = srv.submit([(n, 1, 1)], [(b, n, n)], "f", 0, inp=xs[i], out=outs[i])
                assert rid >= 0
                assert srv.wait(rid)[0] == 0
        except Exception as e:  # surfaced below
            errors.append(e)

    ts = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    assert srv.stats()["completed"] == len(xs)
    for a, c in zip(ref, outs):
        assert torch.equal(a.view(torch.float32), c.view(torch.float32))
    srv.destroy()
