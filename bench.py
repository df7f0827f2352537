#!/usr/bin/env python3
"""Benchmark of the MI355X FFT library on its headline workloads.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Primary workload (the `value` of the JSON line): the large single transform --
1-D C2C complex128, N = 2^24, batch 1, in place, one forward + one backward
transform per step (the four-step path), inputs i.i.d. normal(0,1) from seed 1
already resident in HBM.  GFLOP/s counts 5 N log2 N per transform.

Also measured in the same run and reported in the same JSON line:
  * batched FP32 C2C, N = 1024, batch 4096, out of place (single LDS pass),
    with input buffers rotated over 512 MiB so the 256 MiB Infinity Cache does
    not serve them, plus the cache-warm figure;
  * the mixed-radix large transform, complex128 N = 2^20 * 3 round trip;
  * 3-D R2C float32 512^3 with its error against float64;
  * the uneven batch: 20,000 complex64 items of mixed length (7-smooth and
    prime) with batch 1..64 each, in one persistent launch per tile class;
  * the precision workload: batched complex64 N = 4096, batch 16384, lines
    of 8 tones + noise, plain FP32 against the matrix-core forms (FP32, bf16
    hi+lo, scaled fp16 hi+lo operands), time and error against complex128;
  * serving: open-loop Poisson arrivals (2k, 10k, 50k requests/s) of a mix
    of 1-D C2C, 2-D R2C and 3-D C2C requests through the serving runtime,
    p50 / p99 latency and transforms/s;
  * roofline of the headline: algorithmic bytes of one transform (every
    element read once and written once, 2 x 16 B x 2^24) over the average
    duration of one transform (all its launches) measured with HIP events on
    the launch stream, plus each pass's own bandwidth;
  * HBM traffic per launch from two rocprofv3 --pmc runs of this file
    (FETCH_SIZE, then WRITE_SIZE), scaled by the same counters' reading for a
    plain copy of known size -- child processes started before the GPU is
    used (--no-traffic skips them);
  * the reference's CPU path: its own benchmark program (aocl_fftz_bench,
    built from its sources with OpenMP) on this host at 1 thread and at the
    host's thread count, and the relative L2 difference between the GPU and
    the reference library on the same input (--no-cpu skips these).

One GPU only: --gpus above 1 exits with a message.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "aocl-fftz_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


OUR_KERNELS = ("k_fixed_pass", "k_fft_pass", "k_bluestein", "k_batch", "k_mfma_pass", "k_persist")


def _is_ours(name):
    return any(k in name for k in OUR_KERNELS)


def traffic_probe(manifest_path):
    """Run under `rocprofv3 --pmc <counter>` by measure_traffic(): a plain
    256 MiB device copy (the calibration kernel of known size), then two
    forward and two backward executes of the headline 2^24 complex128
    transform, then three executes of the batched N=1024 x 4096 complex64
    plan.  Writes the dispatch manifest (which of this library's dispatches
    belong to which workload) as JSON."""
    import torch
    import fftz
    assert torch.cuda.is_available()
    fftz.load()
    mib = 256
    a = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    man = {"copy_bytes": mib << 20, "workloads": []}
    n = 1 << 24
    x = torch.randn(n, dtype=torch.complex128, device="cuda")
    fwd = fftz.Plan([(n, 1, 1)], [(1, n, n)], "d", 0, inplace=True, inp=x, out=x)
    bwd = fftz.Plan([(n, 1, 1)], [(1, n, n)], "d", 1, inplace=True, inp=x, out=x)
    nl = fwd.launches()
    for _ in range(2):
        assert fwd.execute() == 0
    for _ in range(2):
        assert bwd.execute() == 0
    torch.cuda.synchronize()
    man["workloads"].append({"name": "2p24", "launches": nl, "executes": 4})
    fwd.destroy()
    bwd.destroy()
    del x
    xs = torch.rand(4096, 1024, dtype=torch.complex64, device="cuda")
    ys = torch.empty_like(xs)
    p = fftz.Plan([(1024, 1, 1)], [(4096, 1024, 1024)], "f", 0)
    for _ in range(3):
        assert p.execute_io(xs, ys) == 0
    torch.cuda.synchronize()
    man["workloads"].append({"name": "n1024", "launches": p.launches(), "executes": 3})
    p.destroy()
    with open(manifest_path, "w") as f:
        json.dump(man, f)


def _counter_rows(outdir, counter):
    """(dispatch id, kernel name, value) for one counter from rocprofv3's
    counter_collection CSV under outdir, one row per dispatch (instances
    summed)."""
    import csv
    import glob
    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    per = {}
    for fn in files:
        with open(fn, newline="") as f:
            for row in csv.DictReader(f):
                r = {k.lower(): v for k, v in row.items() if k}
                if r.get("counter_name", "") != counter:
                    continue
                key = (int(r.get("process_id", 0) or 0), int(r["dispatch_id"]))
                name = r.get("kernel_name", "")
                v = float(r["counter_value"])
                if key in per:
                    per[key] = (name, per[key][1] + v)
                else:
                    per[key] = (name, v)
    return [(k[1], v[0], v[1]) for k, v in sorted(per.items())]


def measure_traffic(budget_s=90):
    """HBM bytes per launch of the headline transform and of the batched
    N=1024 plan, from two rocprofv3 --pmc runs of traffic_probe (FETCH_SIZE,
    then WRITE_SIZE, each in KiB per dispatch), each scaled by what the same
    counter reads for the plain copy kernel of known size.  Runs before this
    process touches the GPU.  Returns (result dict or None, note)."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None, "rocprofv3 not found"
    work = tempfile.mkdtemp(prefix="fftz_traffic_")
    vals = {}
    man = None
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        outdir = os.path.join(work, counter)
        mpath = os.path.join(work, counter + ".json")
        cmd = ["timeout", "-s", "KILL", str(budget_s), prof, "--pmc", counter, "-d", outdir, "-o", "run",
               "--output-format", "csv", "--", sys.executable, os.path.abspath(__file__),
               "--traffic-probe", mpath]
        log(f"[bench] traffic: rocprofv3 --pmc {counter}")
        t0 = time.perf_counter()
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        log(f"[bench] traffic: {counter} pass rc={r.returncode} in {time.perf_counter() - t0:.1f} s")
        if not os.path.exists(mpath):
            return None, f"{counter} probe failed (rc {r.returncode}): {r.stdout[-400:]}"
        with open(mpath) as f:
            man = json.load(f)
        rows = _counter_rows(outdir, counter)
        copies = [v for _, name, v in rows if "copyBuffer" in name]
        ours = [v for _, name, v in rows if _is_ours(name)]
        if not copies or not ours:
            return None, f"{counter}: no copy or FFT dispatches in the counter CSV"
        # the 256 MiB copy is the largest copy dispatch
        factor = man["copy_bytes"] / (max(copies) * 1024.0)
        vals[counter] = (factor, [v * 1024.0 * factor for v in ours])
    res = {"copy_factor": {c: round(vals[c][0], 4) for c in vals}}
    i = 0
    for w in man["workloads"]:
        k, e = w["launches"], w["executes"]
        per_launch = []
        for j in range(k):
            fetch = [vals["FETCH_SIZE"][1][i + x * k + j] for x in range(e)]
            write = [vals["WRITE_SIZE"][1][i + x * k + j] for x in range(e)]
            per_launch.append({"read": sum(fetch) / e, "write": sum(write) / e})
        i += k * e
        res[w["name"]] = per_launch
    if i != len(vals["FETCH_SIZE"][1]) or i != len(vals["WRITE_SIZE"][1]):
        return None, f"dispatch count mismatch: manifest {i}, counters {len(vals['FETCH_SIZE'][1])}"
    return res, "measured in this run"


def launch_bytes(plan_json):
    """Algorithmic HBM bytes of each launch: every element read once and
    written once."""
    out = []
    pb = plan_json["prec"]
    for ln in plan_json["launches"]:
        in_e = pb if ln["flags"] & 512 else 2 * pb
        out_e = pb if ln["flags"] & 1024 else 2 * pb
        nin = 2 * ln["L"] if ln["flags"] & 512 else ln["nvalid_in"]
        nout = ln["L"] + 1 if ln["flags"] & 128 else ln["nvalid_out"]
        if ln["flags"] & 1024:
            nout = 2 * ln["L"]
        out.append(ln["nlines"] * (nin * in_e + nout * out_e))
    return out


def kernel_name(launch, real):
    """Kernel symbol (as rocprofv3 abbreviates it) of one plan launch."""
    if launch["fixed"]:
        return f"k_fixed_pass<{real}, {launch['L']}>"
    kind = "k_bluestein" if launch["kind"] == "bluestein" else "k_fft_pass"
    return f"{kind}<{real}, E={launch['E']}, {launch['block']}>"


def roofline_from_profile(prof, bytes_per_launch):
    """Average duration per launch index; the dominant kernel is the launch
    index with the largest total time."""
    import numpy as np
    arr = np.array(prof, dtype=np.float64)  # executes x launches (ms)
    avg = arr.mean(axis=0)
    k = int(np.argmax(avg))
    achieved = bytes_per_launch[k] / (avg[k] * 1e-3)
    return k, avg, achieved


def bench_roundtrip_2p24(torch, fftz, steps, warmup, n=1 << 24):
    """Large 1-D: N=2^24 complex128 in place, forward + backward per step."""
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    x = torch.randn(n, dtype=torch.complex128, device="cuda", generator=g)
    x0 = x.clone()
    stream = torch.cuda.current_stream()
    fwd = fftz.Plan([(n, 1, 1)], [(1, n, n)], "d", 0, inplace=True, inp=x, out=x)
    bwd = fftz.Plan([(n, 1, 1)], [(1, n, n)], "d", 1, inplace=True, inp=x, out=x)
    for p in (fwd, bwd):
        assert p.set_stream(stream) == 0
    pj = fftz.plan_json([(n, 1, 1)], [(1, n, n)], "d", 0, inplace=True)
    log(f"[bench] 2^24 fp64 plan: {fwd.info}")
    for _ in range(warmup):
        fwd.execute()
        bwd.execute()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    fwd.profile(True)
    bwd.profile(True)
    torch.cuda.synchronize()
    t0.record(stream)
    for _ in range(steps):
        assert fwd.execute() == 0
        assert bwd.execute() == 0
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / steps
    prof_f = fwd.profile_read()
    prof_b = bwd.profile_read()
    fwd.profile(False)
    bwd.profile(False)
    # round trip: x was scaled by N per step; N = 2^24 so the rescale is exact
    k = steps + warmup
    back = x / (float(n) ** k)
    rt_err = (torch.linalg.vector_norm(back - x0) / torch.linalg.vector_norm(x0)).item()
    bpl = launch_bytes(pj)
    kidx, avg, achieved = roofline_from_profile(prof_f + prof_b, bpl)
    import numpy as np
    # one transform = all launches of one execute: the sum of its launch
    # durations (events bracket every launch on the launch stream)
    per_transform = np.array([sum(e) for e in prof_f + prof_b], dtype=np.float64)
    flops = 2 * 5.0 * n * math.log2(n)
    res = {
        "ms_per_step": ms,
        "gflops": flops / (ms * 1e-3) / 1e9,
        "transforms_per_s": 2.0 / (ms * 1e-3),
        "roundtrip_rel_l2": rt_err,
        "launch_ms_avg": [round(v, 4) for v in avg.tolist()],
        "launch_bytes": bpl,
        "launch_kernels": [kernel_name(ln, "double") for ln in pj["launches"]],
        "dominant_launch": kidx,
        "dominant_achieved_Bps": achieved,
        "dominant_kernel": kernel_name(pj["launches"][kidx], "double"),
        "transform_ms_avg": float(per_transform.mean()),
        "transform_bytes": 2 * 16.0 * n,
        "transform_frac_of_hbm": (2 * 16.0 * n * 2) / (ms * 1e-3) / HBM_PEAK,
        "plan": fwd.info,
    }
    # keep the forward input/output of one fresh transform for the reference
    # comparison in the CPU-baseline leg
    xin = x0.clone()
    fwd2 = fftz.Plan([(n, 1, 1)], [(1, n, n)], "d", 0, inplace=True, inp=xin, out=xin)
    fwd2.set_stream(stream)
    fwd2.execute()
    torch.cuda.synchronize()
    res["_x0"] = x0.cpu().numpy()
    res["_y"] = xin.cpu().numpy()
    for p in (fwd, bwd, fwd2):
        p.destroy()
    del x, x0, xin, back
    torch.cuda.empty_cache()
    return res


def bench_batched_n1024(torch, fftz, steps, warmup, n=1024, b=4096):
    """Batched: N=1024 complex64, batch 4096, out of place."""
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    nbuf = 8  # 8 x (32 + 32) MiB = 512 MiB > Infinity Cache
    xs = [(torch.rand(b, n, dtype=torch.complex64, device="cuda", generator=g) * 2 - (1 + 1j)) for _ in range(nbuf)]
    ys = [torch.empty_like(xs[0]) for _ in range(nbuf)]
    stream = torch.cuda.current_stream()
    p = fftz.Plan([(n, 1, 1)], [(b, n, n)], "f", 0)
    p.set_stream(stream)
    pj = fftz.plan_json([(n, 1, 1)], [(b, n, n)], "f", 0)
    bpl = launch_bytes(pj)
    log(f"[bench] n1024 fp32 plan: {p.info}")
    p.execute_many([xs[i % nbuf] for i in range(warmup * nbuf)], [ys[i % nbuf] for i in range(warmup * nbuf)])
    torch.cuda.synchronize()
    reps = max(steps, 1) * 20
    ins = [xs[i % nbuf] for i in range(reps)]
    outs = [ys[i % nbuf] for i in range(reps)]
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    # one C call enqueues all executes, so the host never starves the GPU
    t0.record(stream)
    assert p.execute_many(ins, outs) == 0
    t1.record(stream)
    torch.cuda.synchronize()
    ms_cold = t0.elapsed_time(t1) / reps
    # kernel durations with events around each launch (separate pass)
    p.profile(True)
    assert p.execute_many(ins, outs) == 0
    torch.cuda.synchronize()
    prof = p.profile_read()
    p.profile(False)
    # cache-warm: the same buffers every time
    t0.record(stream)
    p.execute_many([xs[0]] * reps, [ys[0]] * reps)
    t1.record(stream)
    torch.cuda.synchronize()
    ms_warm = t0.elapsed_time(t1) / reps
    # the same bytes moved by a plain device copy, same buffer rotation and
    # back-to-back timing: the practical ceiling for a 32 MiB -> 32 MiB launch
    for i in range(nbuf):
        ys[i].copy_(xs[i])
    torch.cuda.synchronize()
    t0.record(stream)
    for i in range(reps):
        ys[i % nbuf].copy_(xs[i % nbuf])
    t1.record(stream)
    torch.cuda.synchronize()
    ms_copy = t0.elapsed_time(t1) / reps
    _, avg, achieved = roofline_from_profile(prof, bpl)
    flops = 5.0 * n * math.log2(n) * b
    byts = 2 * 8.0 * n * b
    res = {
        "ms_per_execute": ms_cold,
        "gflops": flops / (ms_cold * 1e-3) / 1e9,
        "transforms_per_s": b / (ms_cold * 1e-3),
        "frac_of_hbm": byts / (ms_cold * 1e-3) / HBM_PEAK,
        "kernel_ms_avg": float(avg[0]),
        "kernel_achieved_Bps": achieved,
        "ms_per_execute_cache_warm": ms_warm,
        "gflops_cache_warm": flops / (ms_warm * 1e-3) / 1e9,
        "ms_copy_same_bytes": ms_copy,
        "copy_same_bytes_frac_of_hbm": byts / (ms_copy * 1e-3) / HBM_PEAK,
        "frac_of_copy_same_bytes": ms_copy / ms_cold,
        "plan": p.info,
    }
    p.execute_io(xs[0], ys[0])
    torch.cuda.synchronize()
    res["_x0"] = xs[0].cpu().numpy()
    res["_y"] = ys[0].cpu().numpy()
    p.destroy()
    del xs, ys
    torch.cuda.empty_cache()
    return res


def bench_mixed_radix(torch, fftz, steps, warmup, n=3 << 20):
    """Mixed radix: 1-D C2C complex128 N = 2^20 * 3, in place, forward +
    backward per step; round-trip error after the 1/N scale."""
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    x = torch.randn(n, dtype=torch.complex128, device="cuda", generator=g)
    x0 = x.clone()
    stream = torch.cuda.current_stream()
    fwd = fftz.Plan([(n, 1, 1)], [(1, n, n)], "d", 0, inplace=True, inp=x, out=x)
    bwd = fftz.Plan([(n, 1, 1)], [(1, n, n)], "d", 1, inplace=True, inp=x, out=x)
    for p in (fwd, bwd):
        assert p.set_stream(stream) == 0
    for _ in range(max(warmup, 1)):
        fwd.execute()
        bwd.execute()
        x.mul_(1.0 / n)
    x.copy_(x0)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(steps):
        fwd.execute()
        bwd.execute()
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / steps
    # one more clean round trip for the error
    x.copy_(x0)
    fwd.execute()
    bwd.execute()
    torch.cuda.synchronize()
    rt = (torch.linalg.vector_norm(x / n - x0) / torch.linalg.vector_norm(x0)).item()
    res = {"n": n, "ms_per_step": round(ms, 4), "gflops": round(2 * 5.0 * n * math.log2(n) / (ms * 1e-3) / 1e9, 1),
           "transform_frac_of_hbm": round(2 * 2 * 16.0 * n / (ms * 1e-3) / HBM_PEAK, 4),
           "roundtrip_rel_l2": rt, "plan": fwd.info}
    fwd.destroy()
    bwd.destroy()
    del x, x0
    torch.cuda.empty_cache()
    return res


def bench_r2c_512(torch, fftz, steps, warmup, n=512):
    """3-D R2C float32 512^3 -> complex64 512 x 512 x 257 (real-packed row pass
    + two column passes); error against direct DFT sums at sampled bins
    (float64 torch reductions) and against the library's float64 transform
    of the same input."""
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    x = torch.randn((n, n, n), dtype=torch.float32, device="cuda", generator=g)
    h = n // 2 + 1
    y = torch.empty((n, n, h), dtype=torch.complex64, device="cuda")
    dims = [(n, 1, 1), (n, n, h), (n, n * n, n * h)]
    vecs = [(1, n ** 3, n * n * h)]
    p = fftz.Plan(dims, vecs, "f", 0, fft_type=1, inp=x, out=y)
    stream = torch.cuda.current_stream()
    p.set_stream(stream)
    for _ in range(max(warmup, 1)):
        p.execute_io(x, y)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(steps):
        p.execute_io(x, y)
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / steps
    info = p.info
    p.destroy()
    xd = x.double()
    yd = torch.empty((n, n, h), dtype=torch.complex128, device="cuda")
    q = fftz.Plan(dims, vecs, "d", 0, fft_type=1, inp=xd, out=yd)
    q.set_stream(stream)
    q.execute_io(xd, yd)
    torch.cuda.synchronize()
    q.destroy()
    err = (torch.linalg.vector_norm(y.to(torch.complex128) - yd) / torch.linalg.vector_norm(yd)).item()
    # the definition at sampled bins, X[k2,k1,k0] = sum x[j2,j1,j0] w^(...), as
    # float64 torch reductions (independent of this library), phases reduced
    # exactly in integers; error relative to the spectrum's RMS
    idx = torch.arange(n, device="cuda", dtype=torch.int64)

    def ph(k):
        return torch.polar(torch.ones(n, device="cuda", dtype=torch.float64),
                           (-2 * math.pi / n) * ((idx * k) % n).double())

    bins = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (256, 256, 256), (3, 100, 200), (129, 7, 511),
            (13, 400, 100)]
    rms = (torch.linalg.vector_norm(yd) / math.sqrt(yd.numel())).item()
    worst = 0.0
    for k0, k1, k2 in bins:
        v = (xd.reshape(-1, n).to(torch.complex128) @ ph(k0)).reshape(n, n)
        ref = ((v @ ph(k1)) @ ph(k2)).item()
        worst = max(worst, abs(complex(y[k2, k1, k0].item()) - ref) / rms)
    byts = 4.0 * n ** 3 + 8.0 * n * n * h
    res = {"ms_per_execute": round(ms, 4), "gflops": round(2.5 * n ** 3 * math.log2(n ** 3) / (ms * 1e-3) / 1e9, 1),
           "algorithmic_frac_of_hbm": round(byts / (ms * 1e-3) / HBM_PEAK, 4), "rel_l2_vs_fp64": err,
           "direct_bins_max_err_over_rms": worst, "plan": info}
    del x, y, xd, yd
    torch.cuda.empty_cache()
    return res


def uneven_parts(torch, fftz, items, x, y, steps):
    """The uneven workload's 7-smooth items and prime (Bluestein) items, each
    planned, autotuned and timed alone on the same buffers: where the batch's
    time goes."""
    import workloads
    smooth = set(workloads.smooth7())
    out = {}
    for name, sel in (("smooth", [it for it in items if it[0] in smooth]),
                      ("bluestein", [it for it in items if it[0] not in smooth])):
        p = fftz.BatchPlan(sel, "f", 0)
        assert p.autotune(x, y, reps=2) == 0
        assert p.execute_io(x, y) == 0
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(steps):
            assert p.execute_io(x, y) == 0
        t1.record()
        torch.cuda.synchronize()
        p.profile(True)
        assert p.execute_io(x, y) == 0
        torch.cuda.synchronize()
        prof = p.profile_read()
        p.profile(False)
        out[name] = {"items": len(sel), "elements": int(sum(n * b for n, b, *_ in sel)),
                     "ms_per_execute": round(t0.elapsed_time(t1) / steps, 4),
                     "class_ms": [round(v, 3) for v in prof[-1]] if prof else None}
        p.destroy()
    return out


def bench_uneven(torch, fftz, steps, warmup, parts=False):
    """Uneven batch: 20,000 complex64 C2C items from seed 3 (70% 7-smooth
    lengths in [16, 16384], 30% primes in [101, 8191] on the Bluestein path),
    batch 1..64 each, packed contiguously, forward, out of place; one
    persistent launch per tile class, replayed as a graph.  parts: also the
    smooth and prime items timed alone (uneven_parts)."""
    import workloads
    items, total = workloads.uneven_batch()
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    x = torch.empty(total, dtype=torch.complex64, device="cuda")
    x.view(torch.float32).uniform_(-1.0, 1.0, generator=g)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    t_plan = time.perf_counter()
    p = fftz.BatchPlan(items, "f", 0)
    t_plan = time.perf_counter() - t_plan
    # tile classes' workgroups per CU chosen by the autotuner on these buffers
    t_tune = time.perf_counter()
    assert p.autotune(x, y, reps=2) == 0
    t_tune = time.perf_counter() - t_tune
    assert p.set_stream(stream) == 0
    log(f"[bench] uneven plan ({t_plan * 1e3:.0f} ms, tuned in {t_tune * 1e3:.0f} ms): {p.info}")
    for _ in range(max(warmup, 1)):
        assert p.execute_io(x, y) == 0
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(steps):
        assert p.execute_io(x, y) == 0
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / steps
    p.profile(True)
    assert p.execute_io(x, y) == 0
    torch.cuda.synchronize()
    prof = p.profile_read()
    p.profile(False)
    flops = workloads.uneven_flops(items)
    byts = 2 * 8.0 * total  # every element read once and written once
    transforms = sum(it[1] for it in items)
    res = {
        "items": len(items),
        "transforms": transforms,
        "elements": total,
        "ms_per_execute": ms,
        "gflops": flops / (ms * 1e-3) / 1e9,
        "transforms_per_s": transforms / (ms * 1e-3),
        "frac_of_hbm": byts / (ms * 1e-3) / HBM_PEAK,
        "class_ms": [round(v, 3) for v in prof[-1]] if prof else None,
        "plan_ms": t_plan * 1e3,
        "autotune_ms": t_tune * 1e3,
        "plan": p.info,
    }
    p.destroy()
    if parts:
        res["parts"] = uneven_parts(torch, fftz, items, x, y, steps)
    del x, y
    torch.cuda.empty_cache()
    return res


def sinusoid_lines(torch, n, b, seed=4, tones=8, noise=1e-3, chunk=512):
    """complex64 (b, n): per line, 8 complex tones with random integer
    frequencies, amplitudes in [0.5, 2) and phases, plus noise*normal(0,1)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.empty((b, n), dtype=torch.complex64, device="cuda")
    j = torch.arange(n, device="cuda", dtype=torch.float64)
    for s in range(0, b, chunk):
        m = min(chunk, b - s)
        f = torch.randint(0, n, (m, tones, 1), device="cuda", generator=g).double()
        a = torch.rand((m, tones, 1), device="cuda", generator=g, dtype=torch.float64) * 1.5 + 0.5
        ph = torch.rand((m, tones, 1), device="cuda", generator=g, dtype=torch.float64) * (2 * math.pi)
        ang = (2 * math.pi / n) * f * j + ph
        v = torch.polar(a.expand_as(ang), ang).sum(dim=1)
        v += noise * torch.complex(torch.randn((m, n), device="cuda", generator=g, dtype=torch.float64),
                                   torch.randn((m, n), device="cuda", generator=g, dtype=torch.float64))
        x[s:s + m] = v.to(torch.complex64)
        del ang, v
    return x


def bench_split(torch, fftz, steps, warmup, n=4096, b=16384):
    """Precision workload: batched complex64 C2C, N=4096, batch 16384, lines of
    8 tones + 1e-3 noise; the plain FP32 pass against the matrix-core forms
    (FP32 operands, bf16 hi+lo, scaled fp16 hi+lo).  Error of each against
    numpy's FP64 transform of 256 sampled lines of the same (upcast) input,
    and over all lines against the library's complex128 transform."""
    x = sinusoid_lines(torch, n, b)
    xd = x.to(torch.complex128)
    yd = torch.empty_like(xd)
    stream = torch.cuda.current_stream()
    pd = fftz.Plan([(n, 1, 1)], [(b, n, n)], "d", 0)
    pd.set_stream(stream)
    assert pd.execute_io(xd, yd) == 0
    torch.cuda.synchronize()
    pd.destroy()
    del xd
    y = torch.empty_like(x)
    import numpy as np
    sample = torch.arange(0, b, b // 256, device="cuda")[:256]
    ref_np = np.fft.fft(x[sample].cpu().numpy().astype(np.complex128), axis=1)
    flops = 5.0 * n * math.log2(n) * b
    byts = 2 * 8.0 * n * b
    names = {0: "plain_fp32", 1: "mfma_fp32", 2: "mfma_bf16x3", 3: "mfma_fp16x2"}
    out = {}
    ydn = torch.linalg.vector_norm(yd).item()
    for mode in (0, 1, 2, 3):
        p = fftz.Plan([(n, 1, 1)], [(b, n, n)], "f", 0)
        if mode:
            assert p.set_compute_mode(mode) == 0
        p.set_stream(stream)
        for _ in range(max(warmup, 1)):
            p.execute_io(x, y)
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        for _ in range(steps):
            assert p.execute_io(x, y) == 0
        t1.record(stream)
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / steps
        err = 0.0
        for s in range(0, b, 2048):
            err += torch.linalg.vector_norm(y[s:s + 2048].to(torch.complex128) - yd[s:s + 2048]).item() ** 2
        # independent FP64 reference on 256 sampled lines: numpy's pocketfft
        # of the same (upcast) input
        ys = y[sample].cpu().numpy().astype(np.complex128)
        es = float(np.linalg.norm(ys - ref_np) / np.linalg.norm(ref_np))
        out[names[mode]] = {"ms": round(ms, 4), "gflops": round(flops / (ms * 1e-3) / 1e9, 1),
                            "frac_of_hbm": round(byts / (ms * 1e-3) / HBM_PEAK, 4),
                            "rel_l2_vs_fp64": float(math.sqrt(err) / ydn),
                            "rel_l2_vs_numpy_fp64_256_lines": es}
        log(f"[bench] split n4096 {names[mode]}: {ms * 1e3:.1f} us rel-L2 {out[names[mode]]['rel_l2_vs_fp64']:.3e}")
        p.destroy()
    del x, y, yd
    torch.cuda.empty_cache()
    return out


def bench_serving(torch, fftz, rates=(2000, 10000, 25000, 50000), seconds=0.4, max_count=4000, inflight=16):
    """Serving: open-loop Poisson arrivals of the serving request mix (1-D C2C,
    2-D R2C, 3-D C2C of seeded non-power-of-two sizes, 10% complex128) through
    one server (16 slots, 16 GiB budget); request buffers are carved from a
    ring in one device arena.  Plans are warmed first; requests are submitted
    at their arrival times by the library's native generator and latency runs
    from the arrival time; p50 / p99 latency and throughput per rate."""
    import numpy as np
    import workloads
    reqs = workloads.serving_requests(count=max_count)
    arena_bytes = 4 << 30
    arena = torch.empty(arena_bytes // 4, dtype=torch.float32, device="cuda").uniform_(-1.0, 1.0)
    base = arena.data_ptr()
    srv = fftz.Server(16 << 30, max_inflight=inflight)
    ring = [0]

    def carve(nbytes):
        nbytes = (nbytes + 255) & ~255
        if ring[0] + nbytes > arena_bytes:
            ring[0] = 0
        off = ring[0]
        ring[0] += nbytes
        return base + off

    def request(r):
        dims, vecs, prec, ftype, nin, nout, in_real = r
        esz = 8 if prec == "f" else 16
        inp = carve(nin * (esz // 2 if in_real else esz))
        out = carve(nout * esz)
        return {"dims": dims, "vecs": vecs, "precision": prec, "fft_type": ftype, "inp": inp, "out": out}

    def submit(r):
        q = request(r)
        return srv.submit(q["dims"], q["vecs"], q["precision"], 0, fft_type=q["fft_type"], inp=q["inp"],
                          out=q["out"])

    seen = set()
    for r in reqs:  # warm: one request per distinct problem builds its plan
        key = (str(r[0]), str(r[1]), r[2], r[3])
        if key not in seen:
            seen.add(key)
            srv.wait(submit(r))
    torch.cuda.synchronize()
    out = {"max_inflight": inflight, "distinct_problems": len(seen), "requests_mean_MiB": round(float(np.mean(
        [(r[4] * ((8 if r[2] == "f" else 16) // (2 if r[6] else 1)) + r[5] * (8 if r[2] == "f" else 16)) / 2 ** 20
         for r in reqs])), 3)}
    rng = np.random.default_rng(5)
    for rate in rates:
        count = int(min(max_count, rate * seconds))
        gaps = rng.exponential(1.0 / rate, count)
        arrivals = np.cumsum(gaps)
        batch = [request(reqs[i]) for i in range(count)]
        srv.reset_stats()
        # open loop from native code (aoclfftz_gpu_server_submit_timed): each
        # request is submitted at its arrival time and its latency counts from
        # that time
        ids, late = srv.submit_timed(batch, arrivals)
        for rid in ids:
            srv.wait(rid)
        st = srv.stats()
        out[f"rate_{rate}"] = {"requests": count, "offered_per_s": rate,
                               "submitted_per_s": round(count / float(arrivals[-1]), 1),
                               "max_submit_late_ms": round(late * 1e3, 4),
                               "p50_ms": round(st["p50_ms"], 4), "p99_ms": round(st["p99_ms"], 4),
                               "max_ms": round(st["max_ms"], 4),
                               "transforms_per_s": round(st["transforms_per_s"], 1),
                               "gflops": round(st["gflops"], 1), "queue_peak": int(st["queue_peak"]),
                               "failed": int(st["failed"])}
        log(f"[bench] serving {rate}/s: p50 {st['p50_ms']:.3f} ms p99 {st['p99_ms']:.3f} ms "
            f"{st['transforms_per_s']:.0f} transforms/s")
    srv.destroy()
    del arena
    torch.cuda.empty_cache()
    return out


def copy_ceiling(torch, mib=256, reps=20):
    """Plain device-to-device copy of a known size: the practical HBM
    ceiling on this box, and the calibration kernel for FETCH/WRITE_SIZE."""
    n = mib * (1 << 20) // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        b.copy_(a)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * mib * (1 << 20) / (ms * 1e-3)


def _ref_bench_time(args, timeout=240):
    """Min execution time (seconds) printed by the reference's own benchmark
    program (oracle/_ref/aocl_fftz_bench: calibrated rounds of >= 100 ms,
    two warm-ups, min over the timed iterations, test/bench/performance.c:36-125)."""
    import re
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "aocl_fftz_bench")
    r = subprocess.run([exe] + args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=timeout)
    m = re.search(r"Min Execution time\s*:\s*([0-9.eE+-]+)\s*(ns|us|ms|s)\b", r.stdout)
    if r.returncode != 0 or not m:
        raise RuntimeError(f"reference bench {' '.join(args)} failed (rc {r.returncode}): {r.stdout[-300:]}")
    return float(m.group(1)) * {"ns": 1e-9, "us": 1e-6, "ms": 1e-3, "s": 1.0}[m.group(2)]


def cpu_baseline():
    """The reference's OpenMP CPU path on this host's cores, through its own
    benchmark program built from its sources (min of calibrated rounds,
    test/bench/performance.c:36-125): the 2^24 complex128 in-place transform and the batched
    N=1024 x 4096 complex64 transform, each at 1 thread and at the host's
    thread count(s).  Runs before this process touches the GPU."""
    exe = os.path.join(ROOT, "oracle", "_ref", "aocl_fftz_bench")
    if not os.path.exists(exe):
        return None, {"cpu_baseline_note": "oracle/_ref/aocl_fftz_bench not built: CPU baseline unavailable"}
    avail = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = sorted({1, avail} | ({omp} if 0 < omp <= avail else set()))
    cfgs = {
        "2p24_fp64_inplace": (["16777216", "-b", "p", "-p", "d", "-o", "3", "-r", "i"], 5.0 * (1 << 24) * 24),
        "n1024_b4096_fp32": (["4096v1024", "-b", "p", "-p", "f", "-o", "3"], 5.0 * 1024 * 10 * 4096),
    }
    runs = {}
    for name, (args, flops) in cfgs.items():
        runs[name] = {}
        for t in threads:
            t0 = time.perf_counter()
            sec = _ref_bench_time(args + ["-n", str(t), "-i", "10"])
            runs[name][str(t)] = {"min_ms": round(sec * 1e3, 4), "gflops": round(flops / sec / 1e9, 2)}
            log(f"[bench] cpu reference {name} threads={t}: {sec * 1e3:.3f} ms "
                f"({flops / sec / 1e9:.1f} GFLOP/s; {time.perf_counter() - t0:.1f} s)")
    head = runs["2p24_fp64_inplace"]
    best_t = max(head, key=lambda k: head[k]["gflops"])
    base = {"value": head[best_t]["gflops"], "unit": "GFLOP/s", "cores": int(best_t), "kind": "reference",
            "sample": (f"reference aocl_fftz_bench (built from its sources, OpenMP) "
                       f"'16777216 -b p -p d -o 3 -r i -n {best_t} -i 10': min of 10 calibrated rounds "
                       f"of the N=2^24 complex128 in-place forward transform, {head[best_t]['min_ms']} ms; "
                       f"1 thread: {head['1']['min_ms']} ms ({head['1']['gflops']} GFLOP/s); "
                       f"host affinity {avail} CPUs, OMP_NUM_THREADS={omp or 'unset'}")}
    return base, {"cpu_reference_runs": runs, "cpu_threads_tried": threads}


def cpu_baseline_agreement(res_2p24, res_1024):
    """CPU-baseline leg, second part: relative L2 between the GPU outputs and the reference library's own
    transform of the same inputs (ctypes into oracle/_ref; not timed)."""
    import numpy as np
    from oracle import reflib
    if not reflib.available():
        return {}
    n = 1 << 24
    buf = np.ascontiguousarray(res_2p24["_x0"]).copy()
    plan = reflib.RefPlan([(n, 1, 1)], [(1, n, n)], "d", -1, buf, buf, threads=min(16, os.cpu_count() or 1))
    plan.execute()
    plan.destroy()
    y = res_2p24["_y"]
    out = {"rel_l2_vs_reference_2p24": float(np.linalg.norm(y - buf) / np.linalg.norm(buf))}
    xb = res_1024["_x0"].astype(np.complex128)
    yb = reflib.c2c(xb[:256], -1, "d", threads=min(16, os.cpu_count() or 1))
    out["rel_l2_vs_reference_n1024"] = float(np.linalg.norm(res_1024["_y"][:256] - yb) / np.linalg.norm(yb))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the reference CPU baseline")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM-traffic runs")
    ap.add_argument("--traffic-probe", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--legs", default="mixed,r2c,uneven,split,serving",
                    help="extra workloads after the primary and N=1024 legs (comma list; '' for none; "
                         "uneven_parts: the uneven batch plus its smooth and prime items timed alone)")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args.traffic_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the transforms here "
              "are single-GPU by design and no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)
    if args.steps + args.warmup > 40:
        sys.exit("steps + warmup must stay <= 40 (unnormalised round trips grow by N per step)")
    # CPU baseline and counter runs are child processes started before this
    # process initialises the GPU
    base, extra = (None, {})
    if not args.no_cpu:
        log("[bench] cpu baseline: the reference's own benchmark program")
        base, extra = cpu_baseline()
    traffic, traffic_note = None, "not measured (--no-traffic)"
    if not args.no_traffic:
        traffic, traffic_note = measure_traffic()
        log(f"[bench] traffic: {traffic_note}")
    import torch
    import fftz
    assert torch.cuda.is_available()
    fftz.load()
    # time on a dedicated (default-priority) stream, not the legacy NULL stream
    torch.cuda.set_stream(torch.cuda.Stream())
    log("[bench] start 2^24 fp64 round trip")
    r24 = bench_roundtrip_2p24(torch, fftz, args.steps, args.warmup)
    log(f"[bench] 2^24: {r24['ms_per_step']:.3f} ms/step {r24['gflops']:.1f} GFLOP/s")
    log("[bench] start N=1024 batch 4096 fp32")
    r10 = bench_batched_n1024(torch, fftz, args.steps, args.warmup)
    log(f"[bench] n1024: {r10['ms_per_execute'] * 1e3:.2f} us {r10['gflops']:.1f} GFLOP/s "
        f"frac {r10['frac_of_hbm']:.3f}")
    legs = set(x for x in args.legs.split(",") if x)
    ru = rs = rv = rm = rr = None
    if "mixed" in legs:
        log("[bench] start 2^20*3 fp64 round trip")
        rm = bench_mixed_radix(torch, fftz, args.steps, args.warmup)
        log(f"[bench] 2^20*3: {rm['ms_per_step']:.3f} ms/step {rm['gflops']:.1f} GFLOP/s rt {rm['roundtrip_rel_l2']:.2e}")
    if "r2c" in legs:
        log("[bench] start 512^3 R2C fp32")
        rr = bench_r2c_512(torch, fftz, args.steps, args.warmup)
        log(f"[bench] r2c 512^3: {rr['ms_per_execute']:.3f} ms {rr['gflops']:.1f} GFLOP/s err {rr['rel_l2_vs_fp64']:.2e}")
    if "uneven" in legs or "uneven_parts" in legs:
        log("[bench] start uneven batch (20,000 items) fp32")
        ru = bench_uneven(torch, fftz, args.steps, args.warmup, parts="uneven_parts" in legs)
        log(f"[bench] uneven: {ru['ms_per_execute']:.3f} ms {ru['gflops']:.1f} GFLOP/s frac {ru['frac_of_hbm']:.3f}")
        for name, r in ru.get("parts", {}).items():
            log(f"[bench] uneven {name} items alone: {r['ms_per_execute']:.3f} ms classes {r['class_ms']}")
    if "split" in legs:
        log("[bench] start split-precision matrix-core vs plain, N=4096 batch 16384")
        rs = bench_split(torch, fftz, args.steps, args.warmup)
    if "serving" in legs:
        log("[bench] start serving (Poisson arrivals)")
        rv = bench_serving(torch, fftz)
    copy_bps = copy_ceiling(torch)
    log(f"[bench] copy ceiling {copy_bps / 1e9:.0f} GB/s")
    if not args.no_cpu:
        extra.update(cpu_baseline_agreement(r24, r10))
    # roofline of the headline: the whole four-step transform is the unit (its
    # launches run back to back on one stream); algorithmic bytes = every
    # element read once and written once = 2 x 16 B x 2^24 per transform
    t_s = r24["transform_ms_avg"] * 1e-3
    achieved = r24["transform_bytes"] / t_s
    passes = []
    for j, (b_j, ms_j) in enumerate(zip(r24["launch_bytes"], r24["launch_ms_avg"])):
        pj = {"launch": j, "kernel": r24["launch_kernels"][j], "avg_us": round(ms_j * 1e3, 2),
              "moved_GBps": round(b_j / (ms_j * 1e-3) / 1e9, 1)}
        if traffic and "2p24" in traffic:
            t = traffic["2p24"][j]
            pj["traffic_read"] = int(t["read"])
            pj["traffic_write"] = int(t["write"])
        passes.append(pj)
    tr = None
    if traffic and "2p24" in traffic:
        tr = int(sum(t["read"] + t["write"] for t in traffic["2p24"]))
    line = {
        "metric": "C2C FFT GFLOP/s (5Nlog2N) + %HBM roofline; MFMA-split vs plain at same error",
        "value": round(r24["gflops"], 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r24["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: i.i.d. normal(0,1) complex128 from seed 1, resident in HBM",
        "config": {"workload": "1-D C2C complex128 N=2^24 batch=1 in-place, forward+backward per step "
                               "(four-step path)", "n": 1 << 24, "batch": 1, "transforms_per_step": 2},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": tr,
                     "kernel": f"one four-step transform ({len(r24['launch_bytes'])} launches of "
                               f"{', '.join(sorted(set(r24['launch_kernels'])))}), "
                               f"event-timed average {r24['transform_ms_avg'] * 1e3:.1f} us",
                     "algorithmic_bytes_per_transform": int(r24["transform_bytes"]),
                     "traffic_source": traffic_note,
                     "passes": passes},
        "cpu_baseline": base,
        "roundtrip_rel_l2": r24["roundtrip_rel_l2"],
        "transform_frac_of_hbm": round(r24["transform_frac_of_hbm"], 4),
        "launch_ms_avg": r24["launch_ms_avg"],
        "batched_n1024_fp32": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r10.items()
                               if not k.startswith("_")},
        "uneven_batch_fp32": ru and {k: (round(v, 4) if isinstance(v, float) else v) for k, v in ru.items()},
        "mixed_radix_fp64_n3x2p20": rm,
        "r2c_fp32_512cubed": rr,
        "split_mfma_n4096_b16384": rs,
        "serving": rv,
    }
    if traffic and "n1024" in traffic:
        t = traffic["n1024"][0]
        line["batched_n1024_fp32"]["traffic"] = int(t["read"] + t["write"])
    line["copy_ceiling_GBps"] = round(copy_bps / 1e9, 1)
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
