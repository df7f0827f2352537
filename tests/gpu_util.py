"""Helpers for GPU parity tests: run the library through its C ABI on torch
device buffers and compare against the oracle.

Tolerances (relative L2 against the FP64 oracle on the same seeded input):
  FP32: 2e-7 * log2(N)      FP64: 1e-14 * log2(N)
"""
import math

import numpy as np

import fftz


def tol(n: int, prec: str) -> float:
    lg = max(1.0, math.log2(max(n, 2)))
    return (2e-7 if prec == "f" else 1e-14) * lg


def rand_c(shape, seed, prec="f"):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    return x.astype(np.complex64 if prec == "f" else np.complex128)


def to_dev(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def run_c2c(torch, x, direction=0, inplace=False, prec=None, graph=False):
    """Batched 1-D C2C along the last axis of x (2-D array B x N)."""
    if prec is None:
        prec = "f" if x.dtype == np.complex64 else "d"
    b, n = x.shape
    dx = to_dev(torch, x)
    dy = dx if inplace else torch.empty_like(dx)
    p = fftz.Plan([(n, 1, 1)], [(b, n, n)], prec, direction, inplace=inplace, inp=dx, out=dy)
    if graph:
        p.set_graph(True)
    rc = p.execute()
    assert rc == 0, p.info
    torch.cuda.synchronize()
    y = dy.cpu().numpy()
    info = p.info
    p.destroy()
    return y, info


def run_desc(torch, xin_flat, out_len, dims, vecs, prec, direction, inplace=False, fft_type=0,
             out_dtype=None):
    """General descriptor run on flat device buffers; returns flat output."""
    dx = to_dev(torch, xin_flat)
    if inplace:
        dy = dx
    else:
        dt = out_dtype or xin_flat.dtype
        dy = torch.zeros(out_len, dtype=getattr(torch, np.dtype(dt).name), device="cuda")
    p = fftz.Plan(dims, vecs, prec, direction, fft_type=fft_type, inplace=inplace, inp=dx, out=dy)
    rc = p.execute()
    assert rc == 0, p.info
    torch.cuda.synchronize()
    out = dy.cpu().numpy()
    p.destroy()
    return out
