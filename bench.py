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
  * roofline of the dominant kernel: algorithmic bytes per launch over its
    average duration measured with HIP events on its own stream;
  * the reference's CPU path (AOCL-FFTZ built from its sources, OpenMP) timed
    on this host on a bounded sample of the same workload, with the relative
    L2 difference between the GPU and the reference on the same input.

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


def _traffic(kernel):
    """HBM bytes per dispatch of `kernel` measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (calibrated on a plain copy), from the committed summary."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f).get(kernel)
    except OSError:
        return None


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
    flops = 2 * 5.0 * n * math.log2(n)
    res = {
        "ms_per_step": ms,
        "gflops": flops / (ms * 1e-3) / 1e9,
        "transforms_per_s": 2.0 / (ms * 1e-3),
        "roundtrip_rel_l2": rt_err,
        "launch_ms_avg": [round(v, 4) for v in avg.tolist()],
        "launch_bytes": bpl,
        "dominant_launch": kidx,
        "dominant_achieved_Bps": achieved,
        "dominant_kernel": kernel_name(pj["launches"][kidx], "double"),
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
    + two column passes); error against the library's float64 transform of
    the same input."""
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
    byts = 4.0 * n ** 3 + 8.0 * n * n * h
    res = {"ms_per_execute": round(ms, 4), "gflops": round(2.5 * n ** 3 * math.log2(n ** 3) / (ms * 1e-3) / 1e9, 1),
           "algorithmic_frac_of_hbm": round(byts / (ms * 1e-3) / HBM_PEAK, 4), "rel_l2_vs_fp64": err, "plan": info}
    del x, y, xd, yd
    torch.cuda.empty_cache()
    return res


def bench_uneven(torch, fftz, steps, warmup):
    """Uneven batch: 20,000 complex64 C2C items from seed 3 (70% 7-smooth
    lengths in [16, 16384], 30% primes in [101, 8191] on the Bluestein path),
    batch 1..64 each, packed contiguously, forward, out of place; one
    persistent launch per tile class, replayed as a graph."""
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
    (FP32 operands, bf16 hi+lo, scaled fp16 hi+lo).  Error of each against the
    library's own complex128 transform of the same (upcast) input."""
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
        out[names[mode]] = {"ms": round(ms, 4), "gflops": round(flops / (ms * 1e-3) / 1e9, 1),
                            "frac_of_hbm": round(byts / (ms * 1e-3) / HBM_PEAK, 4),
                            "rel_l2_vs_fp64": float(math.sqrt(err) / ydn)}
        log(f"[bench] split n4096 {names[mode]}: {ms * 1e3:.1f} us rel-L2 {out[names[mode]]['rel_l2_vs_fp64']:.3e}")
        p.destroy()
    del x, y, yd
    torch.cuda.empty_cache()
    return out


def bench_serving(torch, fftz, rates=(2000, 10000, 50000), seconds=0.4, max_count=4000, inflight=16):
    """Serving: open-loop Poisson arrivals of the serving request mix (1-D C2C,
    2-D R2C, 3-D C2C, 10% complex128) through one server (16 slots, 16 GiB
    budget); request buffers are carved from a ring in one device arena.
    Plans are warmed first; p50 / p99 latency and throughput per rate."""
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

    def submit(r):
        dims, vecs, prec, ftype, nin, nout, in_real = r
        esz = 8 if prec == "f" else 16
        inp = carve(nin * (esz // 2 if in_real else esz))
        out = carve(nout * esz)
        return srv.submit(dims, vecs, prec, 0, fft_type=ftype, inp=inp, out=out)

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
        srv.reset_stats()
        ids = []
        t0 = time.perf_counter()
        t_next = t0
        for i in range(count):
            t_next += gaps[i]
            while time.perf_counter() < t_next:
                pass
            ids.append(submit(reqs[i]))
        t_sub = time.perf_counter() - t0
        for rid in ids:
            srv.wait(rid)
        st = srv.stats()
        out[f"rate_{rate}"] = {"requests": count, "offered_per_s": rate,
                               "submitted_per_s": round(count / t_sub, 1),
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


def cpu_baseline(res_2p24, res_1024, budget_s):
    """The reference's OpenMP CPU path, timed on this host (bounded sample)."""
    import numpy as np
    from oracle import reflib
    if not reflib.available():
        return None, {}
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    cores = max(1, min(cores, os.cpu_count() or 1))
    n = 1 << 24
    x = np.ascontiguousarray(res_2p24["_x0"])
    buf = x.copy()
    plan = reflib.RefPlan([(n, 1, 1)], [(1, n, n)], "d", -1, buf, buf, threads=cores)
    t_end = time.perf_counter() + budget_s
    times = []
    first = True
    while time.perf_counter() < t_end or len(times) < 2:
        buf[:] = x
        t0 = time.perf_counter()
        plan.execute()
        times.append(time.perf_counter() - t0)
        if first:
            ref_y = buf.copy()
            first = False
    plan.destroy()
    t = min(times)
    gflops = 5.0 * n * math.log2(n) / t / 1e9
    y = res_2p24["_y"]
    rel = float(np.linalg.norm(y - ref_y) / np.linalg.norm(ref_y))
    extra = {"rel_l2_vs_reference_2p24": rel}
    # batched N=1024 vs the reference's FP64 transform of the upcast input
    xb = res_1024["_x0"].astype(np.complex128)
    yb = reflib.c2c(xb[:256], -1, "d", threads=cores)
    extra["rel_l2_vs_reference_n1024"] = float(np.linalg.norm(res_1024["_y"][:256] - yb) / np.linalg.norm(yb))
    t_f, reps = reflib.time_c2c(1024, 4096, "f", cores, budget_s=min(5.0, budget_s / 2))
    extra["cpu_n1024_gflops"] = 5.0 * 1024 * 10 * 4096 / t_f / 1e9
    base = {"value": gflops, "unit": "GFLOP/s", "cores": cores, "kind": "reference",
            "sample": f"{len(times)} forward executes of the N=2^24 complex128 in-place transform "
                      f"(min time {t * 1e3:.1f} ms), AOCL-FFTZ built from its sources, OpenMP {cores} threads; "
                      f"N=1024 x 4096 complex64: {reps} executes"}
    return base, extra


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default="mixed,r2c,uneven,split,serving",
                    help="extra workloads after the primary and N=1024 legs (comma list; '' for none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the transforms here "
              "are single-GPU by design and no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)
    if args.steps + args.warmup > 40:
        sys.exit("steps + warmup must stay <= 40 (unnormalised round trips grow by N per step)")
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
    if "uneven" in legs:
        log("[bench] start uneven batch (20,000 items) fp32")
        ru = bench_uneven(torch, fftz, args.steps, args.warmup)
        log(f"[bench] uneven: {ru['ms_per_execute']:.3f} ms {ru['gflops']:.1f} GFLOP/s frac {ru['frac_of_hbm']:.3f}")
    if "split" in legs:
        log("[bench] start split-precision matrix-core vs plain, N=4096 batch 16384")
        rs = bench_split(torch, fftz, args.steps, args.warmup)
    if "serving" in legs:
        log("[bench] start serving (Poisson arrivals)")
        rv = bench_serving(torch, fftz)
    copy_bps = copy_ceiling(torch)
    log(f"[bench] copy ceiling {copy_bps / 1e9:.0f} GB/s")
    base, extra = (None, {})
    if not args.no_cpu:
        log("[bench] cpu baseline (reference, bounded sample)")
        base, extra = cpu_baseline(r24, r10, args.cpu_budget)
    achieved = r24["dominant_achieved_Bps"]
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
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": _traffic(r24["dominant_kernel"]),
                     "kernel": f"{r24['dominant_kernel']}, launch #{r24['dominant_launch']} of the four-step plan",
                     "algorithmic_bytes_per_launch": r24["launch_bytes"][r24["dominant_launch"]]},
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
    line["copy_ceiling_GBps"] = round(copy_bps / 1e9, 1)
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
