"""CPU emulation of the library's launch list (numpy), for planner checks.

aoclfftz_gpu_plan_json gives every launch's geometry without a GPU.  This
module replays those launches with the semantics of kernels.hip -- tile
gather with the line decomposition l = l0 + ln0*(l1 + ln1*l2), conjugation,
the L-point transform, four-step twiddle, real-packed recombine/separate,
Bluestein -- on flat numpy buffers sized exactly like the caller's and the
work buffer, and asserts every gathered / scattered index is in bounds.
A planner geometry bug therefore shows up here, on the CPU, as an
out-of-bounds index or a wrong result, before any GPU run.
"""
import numpy as np

K_CONJ_IN, K_CONJ_OUT, K_TW4 = 1, 2, 4
K_SCALE, K_R2C_POST, K_C2R_PRE, K_REAL_IN, K_REAL_OUT = 32, 128, 256, 512, 1024
K_REAL_PROMOTE, K_HERM_IN, K_REAL_PART = 2048, 4096, 8192
K_CHIRP_IN, K_SPEC_OUT, K_CHIRP_OUT = 1 << 14, 1 << 15, 1 << 16


def chirp_conj(n):
    """conj(B_k) = exp(-i pi (k^2 mod 2n) / n), k < n."""
    k = np.arange(n, dtype=np.int64)
    return np.exp(-1j * np.pi * ((k * k) % (2 * n)) / n)


def bluestein_spec(n, m):
    """conj(FFT_m(b)) / m with b_k = B_k, b_{m-k} = B_k (k < n), zero elsewhere."""
    b = np.zeros(m, np.complex128)
    bk = np.conj(chirp_conj(n))
    b[:n] = bk
    b[m - np.arange(1, n)] = bk[1:]
    return np.conj(np.fft.fft(b)) / m


class OOB(AssertionError):
    pass


def _line_offsets(ln, nl):
    l = np.arange(nl, dtype=np.int64)
    l0 = l % ln["ln0"]
    r = l // ln["ln0"]
    l1 = r % ln["ln1"]
    l2 = r // ln["ln1"]
    ioff = l0 * ln["ils"][0] + l1 * ln["ils"][1] + l2 * ln["ils"][2]
    ooff = l0 * ln["ols"][0] + l1 * ln["ols"][1] + l2 * ln["ols"][2]
    return ioff, ooff, l0, l1, l2


def _seq_base(ln, key, l0, l1, l2):
    """Sequence index of each line's element 0 (multi-pass Bluestein products)."""
    g = ln.get(key)
    if g is None:
        return np.zeros_like(l0)
    return ln[key + "_base"] + l0 * g[0] + l1 * g[1] + l2 * g[2]


def _check(idx, size, what):
    if idx.size and (idx.min() < 0 or idx.max() >= size):
        raise OOB(f"{what}: index range [{idx.min()}, {idx.max()}] outside buffer of {size}")


def run(plan, bufs):
    """bufs: dict role -> flat numpy array (complex128 or float64 views).
    Roles: 0 in, 1 out, 2 work.  Arrays are modified in place."""
    for ln in plan["launches"]:
        src = bufs[ln["in_role"]]
        dst = bufs[ln["out_role"]]
        L = ln["L"]
        flags = ln["flags"]
        nl = ln["nlines"]
        ioff, ooff, l0, l1, l2 = _line_offsets(ln, nl)
        # launch offsets are bytes in the plan's precision: complex elements
        # are 2 * prec bytes, real elements prec bytes
        prec = plan["prec"]
        in_es = prec * (2 if np.iscomplexobj(src) else 1)
        out_es = prec * (2 if np.iscomplexobj(dst) else 1)
        assert ln["in_off"] % in_es == 0 and ln["out_off"] % out_es == 0
        ioff = ioff + ln["in_off"] // in_es
        ooff = ooff + ln["out_off"] // out_es
        if flags & K_REAL_IN:
            j = np.arange(L, dtype=np.int64)
            ire = ioff[:, None] + (2 * j)[None, :] * ln["ies"]
            iim = ioff[:, None] + (2 * j + 1)[None, :] * ln["ies"]
            _check(ire, src.size, "real load")
            _check(iim, src.size, "real load")
            x = src[ire] + 1j * src[iim]
        elif flags & K_REAL_PROMOTE:
            # full-length R2C: real element e loaded as (x[e], 0)
            nv = ln["nvalid_in"]
            e = np.arange(nv, dtype=np.int64)
            idx = ioff[:, None] + e[None, :] * ln["ies"]
            _check(idx, src.size, "real promote load")
            x = np.zeros((nl, max(L, nv)), np.complex128)
            x[:, :nv] = src[idx]
        elif flags & K_HERM_IN:
            # full-length C2R: bins 0..n/2 stored, bin e > n/2 is conj(bin n-e)
            n = ln["nvalid_in"]
            e = np.arange(n, dtype=np.int64)
            m = np.where(2 * e <= n, e, n - e)
            idx = ioff[:, None] + m[None, :] * ln["ies"]
            _check(idx, src.size, "hermitian load")
            x = np.zeros((nl, max(L, n)), np.complex128)
            v = src[idx]
            x[:, :n] = np.where((2 * e > n)[None, :], np.conj(v), v)
        elif flags & K_CHIRP_IN:
            # x[g] conj(B_g) for sequence index g < nchirp, zero beyond (only
            # the valid elements are read)
            n = ln["nchirp"]
            e = np.arange(L, dtype=np.int64)
            g = _seq_base(ln, "gin", l0, l1, l2)[:, None] + e[None, :] * ln["gin"][3]
            valid = g < n
            idx = ioff[:, None] + e[None, :] * ln["ies"]
            _check(idx[valid], src.size, "chirp load")
            x = np.zeros((nl, L), np.complex128)
            x[valid] = src[idx[valid]]
            if flags & K_CONJ_IN:
                x = np.conj(x)
            x[valid] *= chirp_conj(n)[g[valid]]
        else:
            nv = ln["nvalid_in"]
            e = np.arange(nv, dtype=np.int64)
            idx = ioff[:, None] + e[None, :] * ln["ies"]
            _check(idx, src.size, "load")
            x = np.zeros((nl, max(L, nv)), np.complex128)
            x[:, :nv] = src[idx]
        if flags & K_CONJ_IN and not flags & K_CHIRP_IN:
            x = np.conj(x)
        if ln["kind"] == "bluestein":
            n = ln["nvalid_in"]
            y = np.fft.fft(x[:, :n], axis=1)
            if flags & K_CONJ_OUT:
                y = np.conj(y)
            nvo = ln["nvalid_out"]
            e = np.arange(nvo, dtype=np.int64)
            idx = ooff[:, None] + e[None, :] * ln["oes"]
            _check(idx, dst.size, "bluestein store")
            dst[idx] = y[:, :nvo].real if flags & K_REAL_PART else y[:, :nvo]
            continue
        if flags & K_C2R_PRE:
            X = x[:, :L + 1]
            k = np.arange(L)
            Xm = np.conj(X[:, L - k])
            w = np.exp(2j * np.pi * k / (2 * L))
            Z = (X[:, :L] + Xm) + 1j * ((X[:, :L] - Xm) * w)
            x = np.conj(Z)
        y = np.fft.fft(x[:, :L], axis=1)
        if flags & K_R2C_POST:
            Z = y
            k = np.arange(L + 1)
            Zk = Z[:, k % L]
            Zm = np.conj(Z[:, (L - k) % L])
            w = np.exp(-2j * np.pi * k / (2 * L))
            X = 0.5 * (Zk + Zm) + w * (Zk - Zm) / 2j
            idx = ooff[:, None] + k[None, :] * ln["oes"]
            _check(idx, dst.size, "recombine store")
            dst[idx] = X
            continue
        nvo = ln["nvalid_out"]
        y = y[:, :nvo]
        e = np.arange(nvo, dtype=np.int64)
        if flags & K_TW4:
            N = ln["tw4_n"]
            p = (l0[:, None] * e[None, :])
            assert p.max() < N, "four-step twiddle exponent exceeds its table"
            y = y * np.exp(-2j * np.pi * p / N)
        if flags & (K_SPEC_OUT | K_CHIRP_OUT):
            g = _seq_base(ln, "gout", l0, l1, l2)[:, None] + e[None, :] * ln["gout"][3]
            n = ln["nchirp"]
            if flags & K_SPEC_OUT:
                m = L * ln["gout"][3]  # the sequence length: pass length x outer count
                assert g.max() < m
                y = np.conj(y) * bluestein_spec(n, m)[g]
            else:
                y = np.where(g < n, np.conj(chirp_conj(n))[np.minimum(g, n - 1)] * y, y)
        if flags & K_CONJ_OUT:
            y = np.conj(y)
        if flags & K_CHIRP_OUT:
            valid = g < ln["nchirp"]
            idx = ooff[:, None] + e[None, :] * ln["oes"]
            _check(idx[valid], dst.size, "chirp store")
            dst[idx[valid]] = y[valid]
        elif flags & K_REAL_PART:
            idx = ooff[:, None] + e[None, :] * ln["oes"]
            _check(idx, dst.size, "real-part store")
            dst[idx] = y.real
        elif flags & K_REAL_OUT:
            ire = ooff[:, None] + (2 * e)[None, :] * ln["oes"]
            iim = ooff[:, None] + (2 * e + 1)[None, :] * ln["oes"]
            _check(ire, dst.size, "real store")
            _check(iim, dst.size, "real store")
            dst[ire] = y.real
            dst[iim] = y.imag
        else:
            idx = ooff[:, None] + e[None, :] * ln["oes"]
            _check(idx, dst.size, "store")
            dst[idx] = y
    return bufs


def check_launch_limits(plan, lds_cap=163840):
    """Launch-shape invariants the kernels rely on."""
    for ln in plan["launches"]:
        L, C, TL, E = ln["L"], ln["C"], ln["TL"], ln["E"]
        assert ln["block"] == C * TL <= 1024
        assert C & (C - 1) == 0 and TL & (TL - 1) == 0
        assert TL * E >= L, "threads x elements must cover the line"
        for r in ln["radix"]:
            K = -(-E // r)
            assert TL * K >= L // r, "stage butterflies not covered"
        assert ln["lds"] <= lds_cap
        need = L + (1 if ln["flags"] & K_C2R_PRE else 0)
        assert ln["LP"] >= (need - 1) + ((need - 1) >> 4) + 1
        # (matrix-core passes keep 64 bytes past the tile for their fp16 line scale)
        assert ln["lds"] == C * ln["LP"] * (16 if plan["prec"] == 8 else 8) + (64 if ln.get("mfma") else 0)
        G = ln["grid"]
        vrow = ln.get("vrow", 0)
        if vrow:
            # ragged column tiles: rows of ln0 lines, vrow / C tiles per row;
            # every line is held by exactly one live slot
            n0 = ln["ln0"]
            assert vrow % C == 0 and vrow >= n0 > vrow - C and ln["nlines"] % n0 == 0
            ntiles = (vrow // C) * (ln["nlines"] // n0)
            v = np.arange(ntiles * C, dtype=np.int64)
            row, cc = v // vrow, v % vrow
            live = cc < n0
            lines = (row * n0 + cc)[live]
            assert np.array_equal(np.sort(lines), np.arange(ln["nlines"])), "ragged slots must cover each line once"
        else:
            ntiles = -(-ln["nlines"] // C)
        # one tile per workgroup; the XCD remap must be a permutation
        assert G == ntiles
        tiles = []
        per = G >> 3
        for b in range(G):
            t0 = b
            if ln["flags"] & 64 and b < per * 8:
                t0 = (b & 7) * per + (b >> 3)
            tiles.append(t0)
        assert sorted(tiles) == list(range(ntiles)), "tile remap not a permutation"
        assert ln["nlines"] < 2 ** 31


def run_batch(plan, x, direction=0):
    """Replay an uneven-batch work queue (aoclfftz_gpu_batch_plan_json) on a
    flat complex128 input; returns the output buffer.  Checks that tiles
    cover every transform exactly once and stay inside the buffers."""
    esz = 2 * plan["prec"]
    nin = plan["in_extent_bytes"] // esz
    nout = plan["out_extent_bytes"] // esz
    assert x.size >= nin
    y = np.zeros(nout, np.complex128)
    seen = set()
    sizes = plan["sizes"]
    for cls in plan["classes"]:
        for desc, nl, blu, ioff, ooff, dist in cls["tiles"]:
            s = sizes[desc]
            n = s["nvalid"]
            # an entry covers 1..16 rounds of C lines
            assert nl <= 16 * s["C"] and s["C"] * s["TL"] <= cls["maxt"]
            assert s["C"] * s["LP"] * esz + 256 <= cls["lds"] <= 163840  # 256 B header: tile index
            assert bool(blu) == (s["L"] != n)
            for l in range(nl):
                i0 = ioff + l * dist
                o0 = ooff + l * dist
                _check(np.array([i0, i0 + n - 1]), nin, "batch load")
                _check(np.array([o0, o0 + n - 1]), nout, "batch store")
                assert o0 not in seen, "transform written twice"
                seen.add(o0)
                seg = x[i0:i0 + n]
                y[o0:o0 + n] = np.fft.fft(seg) if direction == 0 else np.fft.ifft(seg) * n
    return y, seen
