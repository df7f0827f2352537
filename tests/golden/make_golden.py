"""Write golden vectors from the reference library compiled from its sources.

    make -f oracle/ref.mk -j8 && python tests/golden/make_golden.py

Each fixture holds a seeded input x, the reference's output y (its own
aoclfftz_setup_d / aoclfftz_setup_f + aoclfftz_execute on host memory), and
metadata (size, batch, precision, direction, tolerance for comparing a FP64
transform against y).  The fixtures are data only; the reference library
itself never ships.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import reflib  # noqa: E402

CASES = [
    # (n, batch, precision, direction, tol vs FP64 oracle)
    (16, 4, "d", 0, 1e-14),
    (101, 2, "d", 0, 1e-14),      # prime: reference Bluestein path
    (1000, 2, "d", 0, 1e-14),     # mixed radix CT
    (1024, 2, "d", 1, 1e-14),     # pow2 iterative, backward
    (4096, 1, "d", 0, 1e-14),
    (3 * 7 * 11 * 13, 1, "d", 0, 1e-14),
    (1024, 4, "f", 0, 2e-6),      # FP32, the batched N=1024 workload's shape
]


def main():
    if not reflib.available():
        raise SystemExit("build the reference library first: make -f oracle/ref.mk -j8")
    index = {}
    for n, b, prec, direction, tol in CASES:
        rng = np.random.default_rng(1000 + n)
        x = rng.uniform(-1, 1, (b, n)) + 1j * rng.uniform(-1, 1, (b, n))
        x = x.astype(np.complex64 if prec == "f" else np.complex128)
        y = reflib.c2c(x, -1 if direction == 0 else 1, prec)
        meta = {"n": n, "batch": b, "precision": prec, "direction": direction, "tol": tol,
                "source": "reference aoclfftz_setup_%s + aoclfftz_execute, compiled by oracle/ref.mk" % prec}
        name = f"ref_{prec}_n{n}_b{b}_{'fwd' if direction == 0 else 'bwd'}.npz"
        path = os.path.join(HERE, name)
        np.savez_compressed(path, x=x, y=y, meta=json.dumps(meta))
        index[name] = hashlib.sha256(open(path, "rb").read()).hexdigest()
    with open(os.path.join(HERE, "index.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)
    print(json.dumps(index, indent=1))


if __name__ == "__main__":
    main()
