"""Real-transform data layouts of the reference's stride rules
(docs/rst/data_layout_conventions.rst:56-200), shared by the CPU emulator
tests and the GPU tests.

For n real points and B batches (h = n/2 + 1 complex bins), each layout gives
the R2C descriptor (dims, vecs) with real-element strides on the real side and
complex-element strides on the complex side; the C2R descriptor swaps the
sides.  Buffer sizes are in real elements (real side) and complex elements
(complex side); in place, one buffer of the larger size in real elements.
"""
import numpy as np


def layout(kind, n, b):
    """R2C (dims, vecs, inplace) for one layout kind."""
    h = n // 2 + 1
    if kind == "row_oop":        # rows of n reals, rows of h bins
        return [(n, 1, 1)], [(b, n, h)], False
    if kind == "row_inplace":    # rows padded to 2h reals (the doc's 4v50: 52 / 26)
        return [(n, 1, 1)], [(b, 2 * h, h)], True
    if kind == "col_oop":        # batches adjacent, elements b apart on both sides
        return [(n, b, b)], [(b, 1, 1)], False
    if kind == "col_inplace":    # elements 2b reals / b bins apart, batches 2 reals / 1 bin apart
        return [(n, 2 * b, b)], [(b, 2, 1)], True
    if kind == "strided_oop":    # element stride 3 reals / 2 bins, gaps between batches
        return [(n, 3, 2)], [(b, 3 * n + 1, 2 * h + 3)], False
    raise ValueError(kind)


KINDS = ["row_oop", "row_inplace", "col_oop", "col_inplace", "strided_oop"]


def c2r_of(dims, vecs):
    """The C2R descriptor reading the R2C output: sides swapped."""
    return [(d[0], d[2], d[1]) for d in dims], [(v[0], v[2], v[1]) for v in vecs]


def extents(dims, vecs, n):
    """(real-side elements, complex-side elements) of an R2C descriptor."""
    h = n // 2 + 1
    real = 1 + (n - 1) * dims[0][1] + sum((v[0] - 1) * v[1] for v in vecs)
    cplx = 1 + (h - 1) * dims[0][2] + sum((v[0] - 1) * v[2] for v in vecs)
    return real, cplx


def real_index(dims, vecs, n):
    """[b, j] -> real-element offset of input element j of batch b."""
    b = np.arange(vecs[0][0])[:, None]
    j = np.arange(n)[None, :]
    return j * dims[0][1] + b * vecs[0][1]


def cplx_index(dims, vecs, n):
    """[b, k] -> complex-element offset of bin k of batch b."""
    b = np.arange(vecs[0][0])[:, None]
    k = np.arange(n // 2 + 1)[None, :]
    return k * dims[0][2] + b * vecs[0][2]
