"""GPU parity of tuned factorisations: radix orders, complex values per thread
and four-step splits chosen through the tuned cache (the text the autotuner
and the FFTW wisdom calls write), and the autotuner's own choice restored
from that text.

The reference picks these by a cost model (radix plan
core/common/pow2_radix_decompose.c:122-198, radix search
selector/complex/selector_ct_dft.c:104-328); here every non-default choice
must give the DFT within the same tolerance as the default plan
(rel-L2 vs the oracle: FP32 2e-7 log2 N, FP64 1e-14 log2 N), and a plan
rebuilt from the exported text must give the tuned plan's result bit for bit.
"""
import numpy as np
import pytest

import fftz
from gpu_util import rand_c, to_dev, tol
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def key(n, b, prec, inplace):
    return f"p{4 if prec == 'f' else 8}:0:0:{int(inplace)}:0|{n},1,1;|{b},{n},{n};"


def run_plan(torch, x, prec, inplace=False):
    b, n = x.shape
    dx = to_dev(torch, x)
    dy = dx if inplace else torch.empty_like(dx)
    p = fftz.Plan([(n, 1, 1)], [(b, n, n)], prec, 0, inplace=inplace, inp=dx, out=dy)
    assert p.execute() == 0, p.info
    torch.cuda.synchronize()
    info = p.info
    p.destroy()
    return dy.cpu().numpy(), info


CASES = [
    # n, batch, precision, in place, tuned fields, plan-info fragment
    (1024, 64, "f", False, "r=1024:8.8.16", "r=8x8x16"),
    (1024, 64, "f", False, "r=1024:4.4.4.16 e=8", "r=4x4x4x16"),
    (4096, 8, "d", False, "r=4096:16.4.4.16", "r=16x4x4x16"),
    (2000, 8, "f", False, "r=2000:5.5.5.16", "r=5x5x5x16"),
    (1 << 20, 1, "d", True, "f=64.128.128 r=128:8.16", "four-step 64 128 128"),
    (1 << 18, 2, "f", False, "f=512.512", "four-step 512 512"),
    (3 << 16, 1, "f", False, "f=12.128.128 r=12:3.4", "four-step 12 128 128"),
]


@pytest.mark.parametrize("n,b,prec,inplace,fields,frag", CASES)
def test_tuned_factorisation(torch_gpu, n, b, prec, inplace, fields, frag):
    torch = torch_gpu
    print(f"tuned factorisation n={n} batch={b} prec={prec} {fields}", flush=True)
    x = rand_c((b, n), n % 97 + 1, prec)
    line = f"{key(n, b, prec, inplace)} 0 0 0 0 {fields}\n"
    fftz.tune_clear()
    try:
        assert fftz.tune_parse(line) == 1
        y, info = run_plan(torch, x, prec, inplace)
        assert frag in info, info
        ref = O.fft(x.astype(np.complex128))
        assert O.rel_l2(y, ref) <= tol(n, prec), info
        text = fftz.tune_text()
        fftz.tune_clear()
        _, info_default = run_plan(torch, x, prec, inplace)
        assert frag not in info_default
        assert fftz.tune_parse(text) == 1
        y2, info2 = run_plan(torch, x, prec, inplace)
        assert info2 == info
        assert np.array_equal(y2.view(np.uint8), y.view(np.uint8))
    finally:
        fftz.tune_clear()


def test_autotuned_choice_restored_from_text(torch_gpu):
    """The autotuner's choice for a four-step problem (tile shape, split,
    radix orders, values per thread) goes into the cache text; after the cache
    is cleared and the text imported, a new plan is the tuned plan and gives
    its result bit for bit, within tolerance of the oracle."""
    torch = torch_gpu
    n, b, prec = 1 << 18, 2, "f"
    x = rand_c((b, n), 5, prec)
    dx = to_dev(torch, x)
    fftz.tune_clear()
    try:
        p = fftz.Plan([(n, 1, 1)], [(b, n, n)], prec, 0)
        assert p.autotune(dx.clone(), torch.empty_like(dx), reps=2) == 0
        tuned_info = p.info
        p.destroy()
        text = fftz.tune_text()
        entry = [ln for ln in text.splitlines() if ln.startswith(key(n, b, prec, False))]
        assert len(entry) == 1, text
        print("autotuned:", entry[0], "|", tuned_info[:200])
        y, info = run_plan(torch, x, prec)
        assert info == tuned_info
        fftz.tune_clear()
        assert fftz.tune_parse(text) == 1
        y2, info2 = run_plan(torch, x, prec)
        assert info2 == tuned_info
        assert np.array_equal(y2.view(np.uint8), y.view(np.uint8))
        assert O.rel_l2(y, O.fft(x.astype(np.complex128))) <= tol(n, prec)
    finally:
        fftz.tune_clear()
