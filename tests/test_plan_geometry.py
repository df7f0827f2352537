"""Planner geometry checked on the CPU (no GPU): every plan's launch list is
replayed by tests/plan_emulator.py with bounds checks on each access and the
result compared with the oracle."""
import numpy as np
import pytest

import fftz
import plan_emulator as emu
from oracle import oracle as O


def rnd(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)


def c2c_case(dims, vecs, prec="f", direction=0, inplace=False):
    plan = fftz.plan_json(dims, vecs, prec, direction, inplace=inplace)
    emu.check_launch_limits(plan)
    n_in = plan["in_extent_bytes"] // (2 * plan["prec"])
    n_out = plan["out_extent_bytes"] // (2 * plan["prec"])
    x = rnd(n_in, 5)
    out = x.copy() if inplace else np.zeros(n_out, np.complex128)
    work = np.zeros(max(1, plan["work_bytes"] // (2 * plan["prec"])), np.complex128)
    bufs = {0: x.copy() if not inplace else out, 1: out, 2: work}
    emu.run(plan, bufs)
    ref = np.zeros(n_out, np.complex128) if not inplace else x.copy()
    O.c2c_strided(x.copy(), ref, dims, vecs, -1 if direction == 0 else 1)
    err = O.rel_l2(bufs[1], ref)
    assert err < 1e-12, (err, plan)
    return plan


@pytest.mark.parametrize("n", [1, 2, 16, 17, 64, 100, 101, 1024, 4096, 8191, 16384, 12288])
@pytest.mark.parametrize("prec", ["f", "d"])
def test_1d_batched(n, prec):
    b = 3
    c2c_case([(n, 1, 1)], [(b, n, n)], prec)


@pytest.mark.parametrize("n,radix", [(6561, [9, 9, 9, 9]), (2187, [9, 9, 9, 3]), (3003, [13, 11, 7, 3]),
                                     (1008, [16, 9, 7]), (4096, [16, 16, 16])])
def test_default_radix_plan(n, radix):
    """Fewest stages from {16, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2}, largest
    radix first: powers of three run as radix-9 stages (3^8 in four stages,
    not eight)."""
    plan = c2c_case([(n, 1, 1)], [(2, n, n)], "f")
    assert plan["launches"][0]["radix"] == radix


@pytest.mark.parametrize("n,prec,M,radix", [(4099, "f", 8448, [16, 16, 11, 3]), (5099, "f", 10240, [16, 16, 10, 4]),
                                            (101, "f", 208, [16, 13]), (1009, "d", 2048, [16, 16, 8])])
def test_bluestein_tile_length_by_stage_cost(n, prec, M, radix):
    """A Bluestein transform in one LDS tile takes, among the 13-smooth
    extended lengths from the smallest (2n - 1 and up, the reference's rule)
    to the next power of two, the one with the least estimated stage work;
    the plan replays to the DFT."""
    plan = c2c_case([(n, 1, 1)], [(2, n, n)], prec)
    ln = plan["launches"][0]
    assert ln["kind"] == "bluestein" and ln["L"] == M and ln["radix"] == radix


def test_bluestein_multi_pass_keeps_smallest_length():
    """Beyond one tile (complex128 M > 8192) the extended length stays the
    reference's smallest 13-smooth M >= 2n - 1 (4099 -> 8232)."""
    plan = c2c_case([(4099, 1, 1)], [(1, 4099, 4099)], "d")
    assert all(ln["kind"] == "pass" for ln in plan["launches"])
    assert plan["work_bytes"] == 2 * 8232 * 16


@pytest.mark.parametrize("n,prec,direction", [(16411, "f", 0), (16411, "f", 1), (8191, "d", 0), (4099, "d", 1),
                                              (65537, "d", 0)])
def test_bluestein_multi_pass(n, prec, direction):
    """Extended length M above one LDS tile: chirp product on the first pass
    of a four-step FFT_M, spectrum product on its last store, a second FFT_M
    whose last store forms the n outputs."""
    plan = c2c_case([(n, 1, 1)], [(2, n, n)], prec, direction)
    assert sum(1 for ln in plan["launches"] if ln["flags"] & emu.K_CHIRP_IN) >= 1
    assert sum(1 for ln in plan["launches"] if ln["flags"] & emu.K_CHIRP_OUT) >= 1


def test_bluestein_multi_pass_strided_inplace():
    # element stride 2 (interleaved with a second batch), in place, backward
    c2c_case([(16411, 2, 2)], [(2, 1, 1)], "f", 1, inplace=True)
    # the Bluestein axis is the outer axis of a 2-D problem
    c2c_case([(5, 1, 1), (16411, 5, 5)], [(2, 5 * 16411, 5 * 16411)], "f")
    c2c_case([(7, 1, 1), (65537, 7, 7)], [(1, 1, 1)], "d", 0, inplace=True)


@pytest.mark.parametrize("dims,vecs,prec", [
    ([(8, 1, 1), (32768, 8, 8)], [(1, 1, 1)], "f"),                      # 2-D, long outer axis
    ([(32768, 1, 1), (8, 32768, 32768)], [(1, 1, 1)], "f"),              # 2-D, long inner axis
    ([(65536, 1, 1), (4, 65536, 65536), (4, 4 * 65536, 4 * 65536)], [(1, 1, 1)], "f"),  # 3-D
    ([(20000, 1, 1), (3, 20000, 20000)], [(1, 1, 1)], "d"),              # FP64 long axis
    ([(1 << 15, 1, 1)], [(3, 1 << 15, 1 << 15), (2, 3 << 15, 3 << 15)], "f"),  # two batch axes
    ([(1 << 14, 1, 1)], [(3, 1 << 14, 1 << 14), (2, 3 << 14, 3 << 14), (2, 6 << 14, 6 << 14)], "d"),
])
def test_long_axis_generality(dims, vecs, prec):
    """An axis above the single-pass limit in an N-D problem, and a long 1-D
    transform with several batch axes: four-step with the other axes as
    batch lines."""
    c2c_case(dims, vecs, prec)
    c2c_case(dims, vecs, prec, direction=1)


@pytest.mark.parametrize("n,prec", [(1 << 15, "f"), (1 << 17, "d"), (1 << 20, "f"), (1 << 20, "d"),
                                    (3 * (1 << 14), "f"), (30030, "f"), (30030, "d"), (16384, "d"),
                                    (1 << 22, "d")])
def test_fourstep(n, prec):
    plan = c2c_case([(n, 1, 1)], [(2, n, n)], prec)
    assert len(plan["launches"]) >= 2


def test_fourstep_cfg2_shapes():
    # the large single transforms: 2^24 and 2^20*3, FP64, in-place, batch 1
    for n in (1 << 24, 3 * (1 << 20)):
        plan = fftz.plan_json([(n, 1, 1)], [(1, n, n)], "d", 0, inplace=True)
        emu.check_launch_limits(plan)
        assert 2 <= len(plan["launches"]) <= 3


def test_fourstep_2_24_emulated():
    n = 1 << 24
    plan = c2c_case([(n, 1, 1)], [(1, n, n)], "d", inplace=True)
    assert len(plan["launches"]) == 3


@pytest.mark.parametrize("direction", [0, 1])
def test_strided_inplace_backward(direction):
    c2c_case([(48, 3, 3)], [(5, 1, 1)], "f", direction, inplace=True)
    c2c_case([(1 << 15, 2, 2)], [(2, 1, 1)], "f", direction, inplace=True)


@pytest.mark.parametrize("shape", [(16, 16), (64, 48), (8, 12, 10), (101, 8), (3, 5, 7, 4), (128, 128, 128)])
def test_multidim(shape):
    dims = []
    stride = 1
    for n in reversed(shape):
        dims.append((n, stride, stride))
        stride *= n
    c2c_case(dims, [(2, stride, stride)], "f")


def test_multi_vec():
    c2c_case([(64, 1, 1)], [(4, 64, 64), (3, 320, 256)], "f")
    c2c_case([(12, 37, 37)], [(37, 1, 1)], "d")


def real_case(n0_shape, prec, batch=2, inplace=False):
    # row-major real input (padded in place), half-complex output
    shape = n0_shape
    n0 = shape[-1]
    h = n0 // 2 + 1
    dims_r2c = []
    s_in = 2 * h if inplace else n0
    s_out = h
    dims_r2c.append((n0, 1, 1))
    for n in reversed(shape[:-1]):
        dims_r2c.append((n, s_in, s_out))
        s_in *= n
        s_out *= n
    vecs = [(batch, s_in, s_out)]
    plan = fftz.plan_json(dims_r2c, vecs, prec, 0, fft_type=1, inplace=inplace)
    emu.check_launch_limits(plan)
    rng = np.random.default_rng(3)
    real_len = plan["in_extent_bytes"] // plan["prec"]
    xr = rng.standard_normal(real_len)
    out = np.zeros(plan["out_extent_bytes"] // (2 * plan["prec"]), np.complex128)
    work = np.zeros(max(1, plan["work_bytes"] // (2 * plan["prec"])), np.complex128)
    emu.run(plan, {0: xr, 1: out, 2: work})
    # reference: per batch, dense real array -> rfftn
    for b in range(batch):
        xb = np.zeros(shape)
        ob = np.zeros(shape[:-1] + (h,), complex)
        for idx in np.ndindex(*shape):
            off = b * vecs[0][1]
            for ax, i in enumerate(reversed(idx)):
                off += i * dims_r2c[ax][1]
            xb[idx] = xr[off]
        for idx in np.ndindex(*(shape[:-1] + (h,))):
            off = b * vecs[0][2]
            for ax, i in enumerate(reversed(idx)):
                off += i * dims_r2c[ax][2]
            ob[idx] = out[off]
        assert O.rel_l2(ob, np.fft.rfftn(xb)) < 1e-12
    # C2R back
    plan_b = fftz.plan_json([(n0, d[2], d[1]) if i == 0 else (d[0], d[2], d[1]) for i, d in enumerate(dims_r2c)],
                            [(batch, vecs[0][2], vecs[0][1])], prec, 1, fft_type=1, inplace=inplace)
    emu.check_launch_limits(plan_b)
    back = np.zeros(plan_b["out_extent_bytes"] // plan_b["prec"])
    work = np.zeros(max(1, plan_b["work_bytes"] // (2 * plan_b["prec"])), np.complex128)
    emu.run(plan_b, {0: out.copy(), 1: back, 2: work})
    nt = int(np.prod(shape))
    for b in range(batch):
        for idx in list(np.ndindex(*shape))[:50]:
            off = b * vecs[0][1]
            for ax, i in enumerate(reversed(idx)):
                off += i * dims_r2c[ax][1]
            assert abs(back[off] / nt - xr[off]) < 1e-10


@pytest.mark.parametrize("shape", [(16,), (1000,), (8, 16), (12, 10, 20)])
def test_real_packed(shape):
    real_case(shape, "f")


def test_real_packed_inplace():
    real_case((6, 32), "d", inplace=True)


def test_cfg4a_plan_shape():
    # 512^3 R2C: one real-packed pass plus two column passes
    n = 512
    h = n // 2 + 1
    dims = [(n, 1, 1), (n, n, h), (n, n * n, n * h)]
    plan = fftz.plan_json(dims, [(1, n ** 3, n * n * h)], "f", 0, fft_type=1)
    emu.check_launch_limits(plan)
    assert len(plan["launches"]) == 3


def _vec_offsets(vecs, side):
    """Element offsets of every batch index (side 1: input strides, 2: output)."""
    offs = [0]
    for v in vecs:
        offs = [o + j * v[side] for o in offs for j in range(v[0])]
    return offs


@pytest.mark.parametrize("n", [15, 17, 64])
@pytest.mark.parametrize("prec", ["f", "d"])
def test_real_full_length_many_batch_axes(n, prec):
    """R2C / C2R with four batch axes that cannot merge: more than three line
    levels, so the outer axes are looped on the host and their offsets are
    counted in real elements on the real side (odd n: the full-length
    promoted / real-part forms; n = 17: Bluestein; n = 64: real-packed)."""
    h = n // 2 + 1
    r, c = n + 5, h + 2  # gaps between batches on both sides
    vecs_r2c = [(2, r, c), (2, 10 * r, 10 * c), (2, 100 * r, 100 * c), (2, 1000 * r, 1000 * c)]
    plan = fftz.plan_json([(n, 1, 1)], vecs_r2c, prec, 0, fft_type=1)
    emu.check_launch_limits(plan)
    assert len(plan["launches"]) >= 2  # host-looped outer axis
    rng = np.random.default_rng(n)
    xr = rng.standard_normal(plan["in_extent_bytes"] // plan["prec"])
    out = np.zeros(plan["out_extent_bytes"] // (2 * plan["prec"]), np.complex128)
    emu.run(plan, {0: xr, 1: out, 2: np.zeros(1, np.complex128)})
    for oi, oo in zip(_vec_offsets(vecs_r2c, 1), _vec_offsets(vecs_r2c, 2)):
        assert O.rel_l2(out[oo:oo + h], np.fft.rfft(xr[oi:oi + n])) < 1e-12
    # C2R: the same tensor read from the complex side
    vecs_c2r = [(v[0], v[2], v[1]) for v in vecs_r2c]
    plan_b = fftz.plan_json([(n, 1, 1)], vecs_c2r, prec, 1, fft_type=1)
    emu.check_launch_limits(plan_b)
    back = np.zeros(plan_b["out_extent_bytes"] // plan_b["prec"])
    assert back.size == xr.size
    emu.run(plan_b, {0: out.copy(), 1: back, 2: np.zeros(max(1, plan_b["work_bytes"] // 16), np.complex128)})
    for oi in _vec_offsets(vecs_r2c, 1):
        assert np.max(np.abs(back[oi:oi + n] / n - xr[oi:oi + n])) < 1e-10


@pytest.mark.parametrize("kind", ["row_oop", "row_inplace", "col_oop", "col_inplace", "strided_oop"])
@pytest.mark.parametrize("n", [50, 64, 15, 17, 1024])
def test_real_layouts_emulated(kind, n):
    """R2C then C2R in the reference's row / column-major, in / out-of-place
    layouts (docs/rst/data_layout_conventions.rst:56-200)."""
    import real_layouts as RL
    b = 4
    dims, vecs, inplace = RL.layout(kind, n, b)
    nr, nc = RL.extents(dims, vecs, n)
    plan = fftz.plan_json(dims, vecs, "d", 0, fft_type=1, inplace=inplace)
    emu.check_launch_limits(plan)
    rng = np.random.default_rng(n)
    size_r = (max(nr, 2 * nc) + 1) // 2 * 2 if inplace else nr
    buf_r = rng.standard_normal(size_r)
    x = buf_r[RL.real_index(dims, vecs, n)].copy()
    if inplace:
        view = buf_r.view(np.complex128)  # the same memory, complex elements
        emu.run(plan, {0: buf_r, 1: view, 2: np.zeros(1, np.complex128)})
        out = view
    else:
        out = np.zeros(nc, np.complex128)
        emu.run(plan, {0: buf_r, 1: out, 2: np.zeros(1, np.complex128)})
    y = out[RL.cplx_index(dims, vecs, n)]
    assert O.rel_l2(y, np.fft.rfft(x, axis=1)) < 1e-12
    # C2R back
    cd, cv = RL.c2r_of(dims, vecs)
    plan_b = fftz.plan_json(cd, cv, "d", 1, fft_type=1, inplace=inplace)
    emu.check_launch_limits(plan_b)
    work = np.zeros(max(1, plan_b["work_bytes"] // 16), np.complex128)
    if inplace:
        emu.run(plan_b, {0: out, 1: buf_r, 2: work})
        back = buf_r
    else:
        back = np.zeros(nr)
        emu.run(plan_b, {0: out.copy(), 1: back, 2: work})
    assert O.rel_l2(back[RL.real_index(dims, vecs, n)] / n, x) < 1e-12


@pytest.mark.parametrize("dims,vecs,prec", [
    # column axis over rows of 257 lines (odd): ragged tiles of 16 / 8 lines
    ([(257, 1, 1), (512, 257, 257)], [(2, 257 * 512, 257 * 512)], "f"),
    ([(257, 1, 1), (512, 257, 257)], [(1, 257 * 512, 257 * 512)], "d"),
    # 3-D with an odd innermost width and a mixed-radix column length
    ([(33, 1, 1), (96, 33, 33), (40, 33 * 96, 33 * 96)], [(1, 33 * 96 * 40, 33 * 96 * 40)], "f"),
])
def test_ragged_column_tiles(dims, vecs, prec):
    """Column passes over a row of lines that the tile width does not divide
    use ragged tiles (each row's last tile partly idle) instead of
    element-strided row tiles; the emulator checks every line is held by one
    live slot and the result is the DFT."""
    plan = c2c_case(dims, vecs, prec)
    assert any(ln["vrow"] for ln in plan["launches"]), plan["launches"]


def test_r2c_512_cubed_column_passes_are_ragged_and_wide():
    """The 3-D R2C 512^3 plan: its two column passes run 16-line tiles (the
    first over rows of 257 half-complex lines, raggedly) on 512-thread fixed
    kernels."""
    n, h = 512, 257
    pj = fftz.plan_json([(n, 1, 1), (n, n, h), (n, n * n, n * h)], [(1, n ** 3, n * n * h)], "f", 0, fft_type=1)
    emu.check_launch_limits(pj)
    cols = pj["launches"][1:]
    assert [ln["C"] for ln in cols] == [16, 16] and all(ln["flags"] & 24 == 24 for ln in cols)
    assert cols[0]["vrow"] == 272 and cols[1]["vrow"] == 0
    assert all(ln["fixed"] and ln["maxt"] == 512 for ln in cols)
