"""GPU parity of R2C / C2R in the reference's real data layouts
(docs/rst/data_layout_conventions.rst:56-200): row-major and column-major,
in place and out of place, plus a general strided layout; even lengths
(real-packed), odd lengths (full-length promoted / Hermitian), primes above
13 (Bluestein).  The in-place column-major layout interleaves the batches'
real inputs with their complex outputs, so a pass that read another tile's
elements after a store would fail here.

Checked against the oracle's R2C / C2R (oracle.r2c / oracle.c2r, pinned in
tests/test_oracle.py) per batch, gathered through the layout's strides."""
import numpy as np
import pytest

import fftz
import real_layouts as RL
from gpu_util import to_dev, tol
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", RL.KINDS)
@pytest.mark.parametrize("n", [50, 64, 15, 17, 1024, 4096])
@pytest.mark.parametrize("prec", ["f", "d"])
def test_real_layout_r2c_c2r(torch_gpu, kind, n, prec):
    torch = torch_gpu
    print(f"real layout {kind} n={n} prec={prec}", flush=True)
    b = 6
    dims, vecs, inplace = RL.layout(kind, n, b)
    nr, nc = RL.extents(dims, vecs, n)
    rt = torch.float32 if prec == "f" else torch.float64
    ct = torch.complex64 if prec == "f" else torch.complex128
    rng = np.random.default_rng(n + len(kind))
    ridx = RL.real_index(dims, vecs, n)
    cidx = RL.cplx_index(dims, vecs, n)
    if inplace:
        size_r = (max(nr, 2 * nc) + 1) // 2 * 2
        buf = torch.from_numpy(rng.standard_normal(size_r)).to(rt).to("cuda")
        x = buf.cpu().numpy().astype(np.float64)[ridx]
        p = fftz.Plan(dims, vecs, prec, 0, fft_type=1, inplace=True, inp=buf, out=buf)
        assert p.execute() == 0
        torch.cuda.synchronize()
        y = buf.view(ct).cpu().numpy()[cidx]
    else:
        xr = torch.from_numpy(rng.standard_normal(nr)).to(rt).to("cuda")
        yc = torch.zeros(nc, dtype=ct, device="cuda")
        x = xr.cpu().numpy().astype(np.float64)[ridx]
        p = fftz.Plan(dims, vecs, prec, 0, fft_type=1, inp=xr, out=yc)
        assert p.execute() == 0
        torch.cuda.synchronize()
        y = yc.cpu().numpy()[cidx]
    p.destroy()
    assert O.rel_l2(y, O.r2c(x)) <= tol(n, prec), kind
    # C2R from the oracle's spectrum placed in the layout
    cd, cv = RL.c2r_of(dims, vecs)
    z = O.r2c(x)
    if inplace:
        host = np.zeros(size_r // 2, np.complex128)
        host[cidx] = z
        buf = torch.from_numpy(host.astype(np.complex64 if prec == "f" else np.complex128)).to("cuda")
        q = fftz.Plan(cd, cv, prec, 1, fft_type=1, inplace=True, inp=buf, out=buf)
        assert q.execute() == 0
        torch.cuda.synchronize()
        back = buf.view(rt).cpu().numpy().astype(np.float64)[ridx]
    else:
        host = np.zeros(nc, np.complex128)
        host[cidx] = z
        zc = torch.from_numpy(host.astype(np.complex64 if prec == "f" else np.complex128)).to("cuda")
        zc0 = zc.clone()
        xo = torch.zeros(nr, dtype=rt, device="cuda")
        q = fftz.Plan(cd, cv, prec, 1, fft_type=1, inp=zc, out=xo)
        assert q.execute() == 0
        torch.cuda.synchronize()
        back = xo.cpu().numpy().astype(np.float64)[ridx]
        assert torch.equal(zc, zc0)  # out-of-place C2R leaves its input alone
    q.destroy()
    assert O.rel_l2(back, O.c2r(z, n)) <= tol(n, prec), kind
