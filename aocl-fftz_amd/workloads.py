"""Seeded synthetic workloads for the benchmark configurations.

uneven_batch(): 1-D C2C items with lengths drawn from a seed -- 70% from the
7-smooth numbers 2^a 3^b 5^c 7^d in [16, 16384], 30% from the primes in
[101, 8191] (the Bluestein path) -- each with a batch count in 1..64, packed
contiguously (input and output offsets are the running element count).
"""
from __future__ import annotations

import numpy as np


def smooth7(lo: int = 16, hi: int = 16384):
    out = []
    a = 1
    while a <= hi:
        b = a
        while b <= hi:
            c = b
            while c <= hi:
                d = c
                while d <= hi:
                    if d >= lo:
                        out.append(d)
                    d *= 7
                c *= 5
            b *= 3
        a *= 2
    return sorted(out)


def primes(lo: int = 101, hi: int = 8191):
    sieve = np.ones(hi + 1, dtype=bool)
    sieve[:2] = False
    for p in range(2, int(hi ** 0.5) + 1):
        if sieve[p]:
            sieve[p * p::p] = False
    return [int(p) for p in np.nonzero(sieve)[0] if p >= lo]


def uneven_batch(count: int = 20000, seed: int = 3, max_batch: int = 64, smooth_frac: float = 0.7):
    """List of (n, batch, in_offset, out_offset) packed contiguously."""
    rng = np.random.default_rng(seed)
    sm = np.array(smooth7())
    pr = np.array(primes())
    items = []
    off = 0
    for _ in range(count):
        n = int(rng.choice(sm)) if rng.random() < smooth_frac else int(rng.choice(pr))
        b = int(rng.integers(1, max_batch + 1))
        items.append((n, b, off, off))
        off += n * b
    return items, off


def uneven_flops(items) -> float:
    return float(sum(5.0 * n * np.log2(n) * b for n, b, *_ in items))


def sinusoids(n: int, batch: int, seed: int = 4, tones: int = 8, noise: float = 1e-3):
    """Complex64 lines of length n: each a sum of `tones` complex exponentials
    with random integer frequencies, amplitudes in [0.5, 2) and phases, plus
    noise * normal(0, 1) in both parts."""
    rng = np.random.default_rng(seed)
    j = np.arange(n)
    out = np.empty((batch, n), dtype=np.complex64)
    for b in range(batch):
        f = rng.integers(0, n, tones)
        a = rng.uniform(0.5, 2.0, tones)
        ph = rng.uniform(0.0, 2 * np.pi, tones)
        x = (a[:, None] * np.exp(1j * (2 * np.pi * np.outer(f, j) / n + ph[:, None]))).sum(axis=0)
        x = x + noise * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
        out[b] = x.astype(np.complex64)
    return out


def _size_menu(rng, lo: int, hi: int, count: int, prime_frac: float = 0.0):
    """count distinct lengths in [lo, hi], log-uniform: 7-smooth lengths, plus
    a prime_frac share of primes (Bluestein), from rng."""
    sm = [v for v in smooth7(lo, hi)]
    pr = [v for v in primes(lo, hi)] if prime_frac > 0 else []
    out = set()
    while len(out) < count:
        target = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        pool = pr if (pr and rng.random() < prime_frac) else sm
        i = int(np.searchsorted(pool, target))
        out.add(int(pool[min(i, len(pool) - 1)]))
    return sorted(out)


def serving_requests(count: int = 4000, seed: int = 11, max_elems: int = 1 << 21):
    """Request mix for the serving loop: 50% 1-D C2C (N in [64, 2^20]), 25% 2-D
    R2C (N x N, N in [128, 4096]), 25% 3-D C2C (N^3, N in [32, 256]); 10%
    complex128; batch in 1..256, reduced so one request holds at most max_elems
    elements.  Lengths come from seeded menus of log-uniform 7-smooth lengths
    (the 1-D menu also holds primes, which run Bluestein), batches from a
    seeded menu of integers, so the number of distinct problems (resident
    plans) stays bounded.  Each entry: (dims, vecs, precision, fft_type,
    in_elems, out_elems, in_real)."""
    rng = np.random.default_rng(seed)
    menu1 = _size_menu(rng, 64, 1 << 20, 24, prime_frac=0.2)
    menu2 = _size_menu(rng, 128, 4096, 10)
    menu3 = _size_menu(rng, 32, 256, 8)
    batches = sorted(set(int(v) for v in np.round(np.exp(rng.uniform(0, np.log(256), 12)))) | {1, 256})
    out = []
    for _ in range(count):
        kind = int(rng.choice(3, p=[0.5, 0.25, 0.25]))
        prec = "d" if rng.random() < 0.1 else "f"
        b = int(rng.choice(batches))
        if kind == 0:
            n = int(rng.choice(menu1))
            b = max(1, min(b, max_elems // n))
            out.append(([(n, 1, 1)], [(b, n, n)], prec, 0, n * b, n * b, False))
        elif kind == 1:
            n = int(rng.choice(menu2))
            h = n // 2 + 1
            b = max(1, min(b, max_elems // (n * n)))
            out.append(([(n, 1, 1), (n, n, h)], [(b, n * n, n * h)], prec, 1, n * n * b, n * h * b, True))
        else:
            n = int(rng.choice(menu3))
            b = max(1, min(b, max_elems // n ** 3))
            out.append(([(n, 1, 1), (n, n, n), (n, n * n, n * n)], [(b, n ** 3, n ** 3)], prec, 0, n ** 3 * b,
                        n ** 3 * b, False))
    return out
