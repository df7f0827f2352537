"""Pin the CPU oracle (oracle/fft_oracle.c) before trusting its expected results.

The reference holds no golden vectors or known-answer files for this path
(its tests compute everything on the fly: naive DFT, property tests,
serial-vs-concurrent equality).  The oracle is therefore pinned by:
  * the naive long-double DFT of test/bench/dft_reference.c:556-636;
  * closed forms (impulse, DC, single tones are exact);
  * Parseval, linearity and time shift (test/bench/accuracy.c:42-688);
  * numpy.fft (pocketfft) as an independent third-party cross-check;
  * the reference library itself, compiled from its sources (oracle/_ref);
  * committed fixtures under tests/golden/ written from that library.
Each check is chosen so that a plausible mistake (wrong sign, dropped
twiddle, transposed index, wrong chirp, wrong scale) fails it; the
"planted" tests below show that for mistakes planted in the oracle's output.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle import reflib

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

SIZES = [1, 2, 3, 5, 7, 8, 11, 12, 13, 16, 17, 20, 48, 60, 64, 97, 101, 128, 210, 256, 343, 512,
         1000, 1024, 2048, 4096, 8191]


def rnd(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("sign", [-1, 1])
def test_oracle_vs_naive_dft(n, sign):
    x = rnd(n, n)
    a = O.fft(x, sign)
    b = O.dft_naive(x, sign)
    assert O.rel_l2(a, b) < 2e-15 * max(1.0, np.log2(n))


@pytest.mark.parametrize("n", [16, 1024, 1000, 4099, 65536, 3 * 2 ** 12])
def test_oracle_vs_numpy(n):
    x = rnd(n, 7)
    assert O.rel_l2(O.fft(x), np.fft.fft(x)) < 2e-15 * np.log2(n)
    assert O.rel_l2(O.fft(x, +1), np.fft.ifft(x) * n) < 2e-15 * np.log2(n)


def test_oracle_many_lengths_in_one_process():
    """Root tables fetched by a transform stay valid through its leaf DFTs
    even after hundreds of other lengths have gone through the table cache
    (a bounded cache that evicted mid-transform gave wrong odd bins for
    n = 10584 = 12 * 7 * 7 * 6 * 3 after ~250 other lengths)."""
    import workloads
    items, _ = workloads.uneven_batch(300, seed=21, max_batch=8)
    O.clear_cache()
    for i, (n, b, _, _) in enumerate(items):
        x = rnd((b, n), i)
        err = O.rel_l2(O.fft(x), np.fft.fft(x, axis=1))
        assert err < 2e-15 * np.log2(n), (i, n, b, err)


@pytest.mark.parametrize("n", [8, 101, 1024, 4096])
def test_closed_forms(n):
    # impulse at 0 -> all ones; DC -> N delta; tone f -> N delta[k - f]
    x = np.zeros(n, complex)
    x[0] = 1
    assert np.max(np.abs(O.fft(x) - 1)) < 1e-14
    x = np.ones(n, complex)
    y = O.fft(x)
    e = np.zeros(n, complex)
    e[0] = n
    assert np.max(np.abs(y - e)) < 1e-12 * n
    for f in (1, n // 3, n - 1):
        j = np.arange(n)
        x = np.exp(2j * np.pi * f * j / n)
        y = O.fft(x)
        e = np.zeros(n, complex)
        e[f] = n
        assert np.max(np.abs(y - e)) < 1e-12 * n


@pytest.mark.parametrize("n", [64, 1000, 4096])
def test_properties(n):
    x = rnd(n, 3)
    y = rnd(n, 4)
    X, Y = O.fft(x), O.fft(y)
    # Parseval
    assert abs(np.vdot(X, X).real / n - np.vdot(x, x).real) < 1e-12 * n
    # linearity
    a, b = 0.3 - 0.2j, -1.7 + 0.5j
    assert O.rel_l2(O.fft(a * x + b * y), a * X + b * Y) < 1e-14 * np.log2(n)
    # time shift by s -> multiply by w^{ks}
    s = 5
    k = np.arange(n)
    assert O.rel_l2(O.fft(np.roll(x, s)), X * np.exp(-2j * np.pi * k * s / n)) < 1e-13
    # round trip
    assert O.rel_l2(O.fft(X, +1) / n, x) < 1e-14 * np.log2(n)


def test_bluestein_length_matches_reference_rule():
    # smallest {2,3,5,7,11,13}-smooth m >= 2n-1 (core/common/bluestein_utils.c:39-70)
    assert O.bluestein_length(101) == 208
    assert O.bluestein_length(8191) == 16384
    assert O.bluestein_length(17) == 33
    for n in (101, 199, 4099, 8191):
        m = O.bluestein_length(n)
        assert m >= 2 * n - 1
        q = m
        for p in (2, 3, 5, 7, 11, 13):
            while q % p == 0:
                q //= p
        assert q == 1
        for c in range(2 * n - 1, m):
            q = c
            for p in (2, 3, 5, 7, 11, 13):
                while q % p == 0:
                    q //= p
            assert q != 1


@pytest.mark.parametrize("n", [2, 4, 16, 50, 512, 1000])
def test_real_packed(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    ref = np.fft.rfft(x)
    assert O.rel_l2(O.r2c_packed(x), ref) < 1e-14 * max(1, np.log2(n))
    assert O.rel_l2(O.r2c(x), ref) < 1e-14 * max(1, np.log2(n))
    assert O.rel_l2(O.c2r(ref, n), x * n) < 1e-14 * max(1, np.log2(n))


def test_nd_strided():
    x = rnd((3, 4, 6, 10), 11)  # batch 3 of a 4x6x10 transform
    y = O.fftn(x, batch_dims=1)
    assert O.rel_l2(y, np.fft.fftn(x, axes=(1, 2, 3))) < 1e-14
    # strided descriptor: every other complex element, batch in column layout
    n, b = 12, 5
    buf = rnd(2 * n * b, 13)
    out = np.zeros_like(buf)
    O.c2c_strided(buf, out, [(n, 2 * b, 2 * b)], [(b, 1, 1)], -1)
    for i in range(b):
        seq = buf[i::2 * b][:n]
        assert O.rel_l2(out[i::2 * b][:n], np.fft.fft(seq)) < 1e-14


# --- error metrics: values fixed by hand ------------------------------------
# orc_rel_l2 is the metric behind every parity assertion, so it is pinned to
# values worked out on paper, on inputs chosen so that the plausible mistakes
# give a different number: a missing square root (5/9 instead of sqrt(5/9)),
# dividing by ||a|| instead of ||b|| (sqrt(5/6)), or a max-norm ratio (2/sqrt(8)).

def test_rel_l2_hand_values():
    a = np.array([1 + 1j, 2 + 0j])
    b = np.array([1 + 0j, 2 + 2j])
    # a - b = (i, -2i): |a-b|^2 = 1 + 4 = 5; |b|^2 = 1 + 8 = 9
    assert O.rel_l2(a, b) == pytest.approx(np.sqrt(5.0 / 9.0), rel=1e-15)
    for wrong in (5.0 / 9.0, np.sqrt(5.0 / 6.0), 2.0 / np.sqrt(8.0)):
        assert abs(O.rel_l2(a, b) - wrong) > 1e-3
    # real arrays: (3, 0) vs (0, 4) -> sqrt(9 + 16) / 4 = 1.25
    assert O.rel_l2(np.array([3.0, 0.0]), np.array([0.0, 4.0])) == pytest.approx(1.25, rel=1e-15)
    # the metric is not symmetric: swapping the operands divides by ||a|| = 3
    assert O.rel_l2(np.array([0.0, 4.0]), np.array([3.0, 0.0])) == pytest.approx(5.0 / 3.0, rel=1e-15)
    # zero reference: the absolute L2 norm of the difference
    assert O.rel_l2(np.array([3.0, 4.0]), np.zeros(2)) == pytest.approx(5.0, rel=1e-15)
    assert O.rel_l2(np.zeros(3), np.zeros(3)) == 0.0
    # equal arrays
    assert O.rel_l2(b, b) == 0.0


def test_rel_l2_sums_every_element():
    # one differing element among many: sqrt(1 / (n * 4)) with b = 2 everywhere;
    # a metric that looked at only the first or last element would miss it
    n = 1001
    b = np.full(n, 2.0 + 0j)
    a = b.copy()
    a[517] += 1.0
    assert O.rel_l2(a, b) == pytest.approx(np.sqrt(1.0 / (4.0 * n)), rel=1e-14)
    a = b.copy()
    a[3] += 1j
    a[900] -= 2.0
    assert O.rel_l2(a, b) == pytest.approx(np.sqrt(5.0 / (4.0 * n)), rel=1e-14)


def test_rel_l2_resolves_tiny_differences():
    # an FP64 ulp-level difference must not vanish (long-double accumulation)
    b = np.ones(4, dtype=np.complex128)
    a = b.copy()
    a[0] = 1.0 + 2.0 ** -52
    # |a - b| = 2^-52 over ||b|| = sqrt(4) = 2 (four complex ones)
    assert O.rel_l2(a, b) == pytest.approx(2.0 ** -52 / 2.0, rel=1e-12)


def test_ref_rel_err_hand_values():
    # compare.c:88-147: max over scalars of |a_i - b_i|, divided by the largest
    # min(|a_i|, |b_i|) over scalars.
    # scalars a = (4, 1), b = (1, 2): errors (3, 1) -> 3; mins (1, 1) -> 1 => 3.
    # Wrong forms give 1.5 (max |b|), 0.75 (max |a|), sqrt(10/5) (L2).
    a = np.array([4.0 + 1.0j])
    b = np.array([1.0 + 2.0j])
    assert O.ref_rel_err(a, b) == pytest.approx(3.0, rel=1e-15)
    # scalars a = (1, -2, 5, 0.5), b = (1.5, -2, 3, 0.5): max err 2; mins
    # (1, 2, 3, 0.5) -> 3 => 2/3
    a = np.array([1 - 2j, 5 + 0.5j])
    b = np.array([1.5 - 2j, 3 + 0.5j])
    assert O.ref_rel_err(a, b) == pytest.approx(2.0 / 3.0, rel=1e-15)
    # sign of the error does not matter, and the real and imaginary parts are
    # separate scalars (not complex moduli): (0 + 3i) vs (0 + 0i) with a
    # companion element of magnitude 4 -> 3 / 4
    a = np.array([0 + 3j, 4 + 0j])
    b = np.array([0 + 0j, 4 + 0j])
    assert O.ref_rel_err(a, b) == pytest.approx(0.75, rel=1e-15)
    # all-zero magnitudes: the absolute error is returned; all zero: 0
    assert O.ref_rel_err(np.array([1.0 + 0j]), np.array([0j])) == pytest.approx(1.0)
    assert O.ref_rel_err(np.zeros(2, complex), np.zeros(2, complex)) == 0.0


# --- planted mistakes: each must be caught by the checks above -------------

def test_planted_sign_error_is_caught():
    n = 64
    j = np.arange(n)
    x = np.exp(2j * np.pi * 3 * j / n)
    wrong = O.fft(x, +1)  # planted: backward used for forward
    e = np.zeros(n, complex)
    e[3] = n
    assert np.max(np.abs(wrong - e)) > 1.0


def test_planted_dropped_twiddle_is_caught():
    # a transform that skips the inter-stage twiddle equals a 2-D DFT of the
    # reshaped data; it must fail against the naive DFT
    n1, n2 = 8, 16
    x = rnd(n1 * n2, 5)
    no_tw = np.fft.fft(np.fft.fft(x.reshape(n1, n2), axis=0), axis=1).T.reshape(-1)
    assert O.rel_l2(no_tw, O.dft_naive(x)) > 0.1


def test_planted_bluestein_chirp_error_is_caught():
    n = 101
    x = rnd(n, 9)
    good = O.fft(x)
    k = np.arange(n)
    wrong = good * np.exp(1j * np.pi * k / n)  # chirp with k instead of k^2
    assert O.rel_l2(wrong, O.dft_naive(x)) > 0.1


# --- reference library (compiled from its sources) --------------------------

needs_ref = pytest.mark.skipif(not reflib.available(), reason="oracle/_ref not built (make -f oracle/ref.mk)")


@needs_ref
@pytest.mark.parametrize("n", [16, 101, 1000, 1024, 4096, 8191, 65536, 3 * 2 ** 12])
def test_oracle_vs_reference_fp64(n):
    x = rnd((2, n), n + 1)
    r = reflib.c2c(x, -1, "d")
    assert O.rel_l2(O.fft(x), r) < 4e-15 * np.log2(n)
    rb = reflib.c2c(x, +1, "d")
    assert O.rel_l2(O.fft(x, +1), rb) < 4e-15 * np.log2(n)


@needs_ref
def test_oracle_vs_reference_fp32_cfg1_shape():
    # the reference's own FP32 transform of the N=1024, B=64 slice
    x = rnd((64, 1024), 1).astype(np.complex64)
    r = reflib.c2c(x, -1, "f")
    o = O.fft(x.astype(np.complex128))
    assert O.rel_l2(r, o) < 2e-7 * 10


def _golden_files():
    if not os.path.isdir(GOLDEN):
        return []
    return sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))


@pytest.mark.parametrize("fname", _golden_files())
def test_oracle_vs_golden(fname):
    d = np.load(os.path.join(GOLDEN, fname))
    meta = json.loads(str(d["meta"]))
    x = d["x"]
    sign = -1 if meta["direction"] == 0 else 1
    y = O.fft(x.astype(np.complex128), sign)
    assert O.rel_l2(y, d["y"]) < meta["tol"]
