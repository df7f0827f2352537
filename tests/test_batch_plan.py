"""Uneven-batch work queue checked on the CPU: every transform of every
item is covered exactly once by the tiles, tiles respect their class's
thread and LDS limits, and replaying the queue reproduces the oracle."""
import numpy as np
import pytest

import fftz
import plan_emulator as emu
import workloads
from oracle import oracle as O


def test_generator_matches_spec():
    items, total = workloads.uneven_batch(2000, seed=3)
    ns = np.array([it[0] for it in items])
    bs = np.array([it[1] for it in items])
    assert bs.min() >= 1 and bs.max() <= 64
    primes = set(workloads.primes())
    frac_prime = np.mean([n in primes for n in ns])
    assert 0.25 < frac_prime < 0.35
    assert all((n in primes) or (16 <= n <= 16384) for n in ns)
    assert sum(n * b for n, b, *_ in items) == total


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("direction", [0, 1])
def test_batch_queue_emulated(prec, direction):
    items, total = workloads.uneven_batch(120, seed=11, max_batch=6)
    if prec == "d":
        items = [it for it in items if it[0] <= 4096 or it[0] in (8192,)]
        off = 0
        packed = []
        for n, b, _, _ in items:
            packed.append((n, b, off, off))
            off += n * b
        items, total = packed, off
    plan = fftz.batch_plan_json(items, prec, direction)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, total) + 1j * rng.uniform(-1, 1, total)
    y, seen = emu.run_batch(plan, x, direction)
    assert len(seen) == sum(b for _, b, *_ in items)
    sign = -1 if direction == 0 else 1
    for n, b, io, oo in items[:40]:
        ref = O.fft(x[io:io + n * b].reshape(b, n), sign).reshape(-1)
        assert O.rel_l2(y[oo:oo + n * b], ref) < 1e-12


@pytest.mark.parametrize("prec", ["f", "d"])
def test_batch_queue_edge_lengths(prec):
    """Length-1 items, one-stage lengths and small Bluestein primes in one
    queue: every transform covered once, in bounds, equal to the DFT."""
    sizes = [(1, 5), (2, 3), (3, 1), (5, 4), (7, 2), (8, 6), (13, 3), (16, 2), (17, 4), (31, 1), (97, 3),
             (1, 1), (101, 2), (1024, 2), (4099, 1), (2, 64), (15, 9)]
    items, off = [], 0
    for n, b in sizes:
        items.append((n, b, off, off))
        off += n * b
    plan = fftz.batch_plan_json(items, prec, 0)
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, off) + 1j * rng.uniform(-1, 1, off)
    y, seen = emu.run_batch(plan, x, 0)
    assert len(seen) == sum(b for _, b in sizes)
    for n, b, io, oo in items:
        ref = O.fft(x[io:io + n * b].reshape(b, n), -1).reshape(-1)
        assert O.rel_l2(y[oo:oo + n * b], ref) < 1e-12


def test_full_size_queue_shape():
    items, total = workloads.uneven_batch(20000, seed=3)
    plan = fftz.batch_plan_json(items, "f", 0)
    ntiles = sum(len(c["tiles"]) for c in plan["classes"])
    lines = sum(t[1] for c in plan["classes"] for t in c["tiles"])
    assert lines == sum(b for _, b, *_ in items)
    assert ntiles >= 20000
    assert plan["in_extent_bytes"] == total * 8
