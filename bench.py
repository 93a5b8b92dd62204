#!/usr/bin/env python3
"""Benchmark of the MI355X PLFFT 1D transform path.

One step = one batched transform pass over data already resident in HBM.
The headline workload is C2C f32 forward, N=4096, batch 32768 (2^27 complex
elements, 1 GiB in + 1 GiB out), out-of-place, unnormalized; further cases of
the same family are reported under "cases". Prints ONE JSON line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "perf-libs-fft_amd"))

HBM_PEAK_GBS = 8000.0


def flops_c2c(n: int, batch: int) -> float:
    return 5.0 * n * math.log2(n) * batch


def settle(torch, plan, x, y, ms: float = 150.0) -> float:
    """Run the transform untimed until `ms` of wall time have passed.

    From idle, an MI355X runs the first ~10 ms of a memory-bound load at
    shifting clocks (a kernel trace of 25 back-to-back 4096-point launches
    read 390 us, rising to 470 us, falling back to 410 us), so a short timed
    region straight after a few warm-up launches measures the power-state
    transition rather than the transform. Returns the settle time spent."""
    t0 = time.perf_counter()
    while True:
        for _ in range(8):
            plan.execute_async(x, y, stream=torch.cuda.current_stream())
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) * 1e3
        if dt >= ms:
            return dt


def time_plan(torch, plan, x, y, steps: int, warmup: int, adjoint: bool = False):
    """Mean ms per launch over `steps` launches, HIP events on the launch stream."""
    stream = torch.cuda.current_stream()
    run = (lambda: plan.adjoint_async(x, y, stream=stream)) if adjoint else \
        (lambda: plan.execute_async(x, y, stream=stream))
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        run()
    end.record(stream)
    end.synchronize()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps


def load_traffic(workload: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summaries."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("traffic_bytes")


def cpu_baseline(n: int, seconds: float = 10.0):
    """The oracle's FP32 port of the reference's CT plan, OpenMP over a
    bounded batch sample of the headline workload, on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle
    from plfft_amd import rng

    threads = min(16, os.cpu_count() or 1)
    sample = 2048
    x = rng.complex_uniform(0, n * sample)
    y = np.empty_like(x)
    oracle.c2c_batched_f32(x, n, -1, sample, 1, n, 1, n, y, threads)  # warm-up
    iters, t0 = 0, time.perf_counter()
    while True:
        oracle.c2c_batched_f32(x, n, -1, sample, 1, n, 1, n, y, threads)
        iters += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    gflops = flops_c2c(n, sample) * iters / dt / 1e9
    return {"value": round(gflops, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"C2C f32 fwd N={n}, {sample} transforms x {iters} passes "
                      f"({dt:.1f} s); CPU restatement, not PLFFT: the oracle's FP32 port "
                      "of the reference CT plan (the AArch64 reference cannot run on this "
                      "x86 host), OpenMP over the batch; its direct codelets and radix-r "
                      "stages are plain O(r^2) loops, so it understates a tuned CPU FFT "
                      "(see cpu_scipy)"}


def cpu_scipy(n: int, seconds: float = 5.0):
    """scipy.fft (pocketfft) on the same sample, all worker threads: a tuned
    CPU FFT beside the port (neither is PLFFT)."""
    import numpy as np
    import scipy.fft

    from plfft_amd import rng

    threads = min(16, os.cpu_count() or 1)
    sample = 2048
    x = rng.complex_uniform(0, n * sample).reshape(sample, n)
    scipy.fft.fft(x, axis=1, workers=threads)
    iters, t0 = 0, time.perf_counter()
    while True:
        scipy.fft.fft(x, axis=1, workers=threads)
        iters += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    return {"value": round(flops_c2c(n, sample) * iters / dt / 1e9, 3), "unit": "GFLOP/s",
            "cores": threads, "kind": "scipy.fft (pocketfft), not PLFFT",
            "sample": f"C2C complex64 fwd N={n}, {sample} transforms x {iters} passes ({dt:.1f} s)"}


def time_stream(torch, fn, steps: int, warmup: int):
    """Mean ms per call of fn(stream) over `steps` calls, HIP events on the
    launch stream."""
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        fn(stream)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn(stream)
    end.record(stream)
    end.synchronize()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps


MIXED_COUNT, MIXED_SEED = 10000, 2024


def progress(msg: str) -> None:
    """Progress lines go to stderr; stdout carries only the JSON line."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bench_mixed(torch, P, steps: int, warmup: int, fused_only: bool = False):
    """The mixed-length batch: 10,000 C2C members, N log-uniform in
    [16, 65536], 10 % moved to a prime, packed back to back, forward,
    out-of-place. Fused = one mixed plan: two tile launches (one per
    dependency level) and three per group of chirp-z members of equal padded
    length; plain = one ordinary plan per member,
    executed one after another
    (eagerly, and captured in a graph to take host launch cost out)."""
    import numpy as np

    from plfft_amd import rng

    lengths = rng.mixed_lengths(MIXED_SEED, MIXED_COUNT)
    total = int(sum(lengths))
    offs = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.int64)
    flops = sum(5.0 * n * math.log2(n) for n in lengths)
    res = {"workload": f"{MIXED_COUNT} C2C members, N log-uniform in [16, 65536], 10% moved "
                       "to a prime, forward, out-of-place, packed",
           "members": MIXED_COUNT, "points": total, "distinct_lengths": len(set(lengths)),
           "primes": int(sum(rng.is_prime(n) for n in lengths))}
    x = torch.empty(total, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    x.view(torch.float32).uniform_(-1.0, 1.0, generator=g)
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    mp = P.MixedPlan(lengths, P.PLFFT_FORWARD, "f32")
    res["plan_seconds"] = round(time.perf_counter() - t0, 2)
    progress(f"mixed plan {res['plan_seconds']} s: {mp.describe()}")
    res["plan"] = mp.describe()
    ws = torch.empty(max(1, mp.workspace_size), dtype=torch.uint8, device="cuda")
    fused = time_stream(torch, lambda s: mp.execute_async(x, y, stream=s, workspace=ws),
                        steps, warmup)
    res["fused_ms"] = round(fused, 4)
    progress(f"mixed fused f32 {fused:.4f} ms")
    res["fused_GFLOP/s"] = round(flops / (fused * 1e-3) / 1e9, 1)
    gbs = 2 * total * 8 / (fused * 1e-3) / 1e9
    res["fused_GB/s"] = round(gbs, 1)
    res["fused_hbm_frac"] = round(gbs / HBM_PEAK_GBS, 4)

    # accuracy of the fused f32 result: numpy FP64 FFT of every 10th member
    torch.cuda.synchronize()
    xh = x.cpu().numpy().astype(np.complex128)
    yh = y.cpu().numpy().astype(np.complex128)
    num = den = 0.0
    worst = 0.0
    for i in range(0, MIXED_COUNT, 10):
        a, n = int(offs[i]), lengths[i]
        ref = np.fft.fft(xh[a:a + n])
        d2 = float(np.sum(np.abs(yh[a:a + n] - ref) ** 2))
        r2 = float(np.sum(np.abs(ref) ** 2))
        num += d2
        den += r2
        worst = max(worst, math.sqrt(d2 / r2))
    res["rel_err"] = float(f"{math.sqrt(num / den):.3g}")
    res["rel_err_max_member"] = float(f"{worst:.3g}")
    res["rel_err_sample"] = "every 10th member vs numpy FP64 FFT"
    if fused_only:  # counter collection: the f32 fused plan only
        return res

    # 16-bit storage, FP32 arithmetic
    for dt in ("f16", "bf16"):
        hp = P.MixedPlan(lengths, P.PLFFT_FORWARD, dt)
        xs = torch.empty(2 * total, dtype=torch.float16 if dt == "f16" else torch.bfloat16,
                         device="cuda")
        xs.uniform_(-1.0, 1.0, generator=g)
        ys = torch.empty_like(xs)
        hws = torch.empty(max(1, hp.workspace_size), dtype=torch.uint8, device="cuda")
        hms = time_stream(torch, lambda s: hp.execute_async(xs, ys, stream=s, workspace=hws),
                          steps, warmup)
        res[f"fused_{dt}_ms"] = round(hms, 4)
        progress(f"mixed fused {dt} {hms:.4f} ms")
        del hp, xs, ys, hws

    # Q0.15: the reference's fixed-point plans take powers of two only, so
    # the lengths are rounded down to powers of two (at most 32768)
    qlen = [min(32768, 1 << (n.bit_length() - 1)) for n in lengths]
    qp = P.MixedPlan(qlen, P.PLFFT_FORWARD, "q15")
    qx = torch.randint(-32768, 32768, (2 * sum(qlen),), dtype=torch.int16, device="cuda")
    qy = torch.empty_like(qx)
    qms = time_stream(torch, lambda s: qp.execute_async(qx, qy, stream=s), steps, warmup)
    res["fused_q15_pow2_ms"] = round(qms, 4)
    res["q15_pow2_points"] = int(sum(qlen))
    res["q15_plan"] = qp.describe()
    progress(f"mixed fused q15 (powers of two) {qms:.4f} ms")
    del qp, qx, qy

    # plain: one plan per distinct length (creation not timed)
    t0 = time.perf_counter()
    plans = {}
    for i, n in enumerate(sorted(set(lengths))):
        plans[n] = P.Plan("c2c", n, P.PLFFT_FORWARD, "f32")
        if i % 1000 == 999:
            progress(f"plain plans: {i + 1} in {time.perf_counter() - t0:.1f} s")
    res["plain_plan_seconds"] = round(time.perf_counter() - t0, 2)
    xp, yp = x.data_ptr(), y.data_ptr()
    jobs = [(plans[n], xp + 8 * int(o), yp + 8 * int(o)) for n, o in zip(lengths, offs)]

    def plain(s):
        for pl, a, b in jobs:
            pl.execute_async(a, b, stream=s)

    psteps = max(2, min(steps, 5))
    res["plain_eager_ms"] = round(time_stream(torch, plain, psteps, 1), 3)
    progress(f"plain eager {res['plain_eager_ms']} ms")
    try:
        gs = torch.cuda.Stream()
        gs.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(gs):
            plain(gs)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            plain(torch.cuda.current_stream())
        torch.cuda.synchronize()
        res["plain_graph_ms"] = round(time_stream(torch, lambda s: graph.replay(), psteps, 1), 3)
        del graph
    except Exception as e:  # reported, not required
        res["plain_graph_ms"] = None
        res["plain_graph_error"] = str(e)[:200]
    res["speedup_vs_plain_eager"] = round(res["plain_eager_ms"] / fused, 2)
    if res.get("plain_graph_ms"):
        res["speedup_vs_plain_graph"] = round(res["plain_graph_ms"] / fused, 2)
    del plans, jobs, mp, ws, x, y
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return res


def case_specs(total: int):
    """The measured cases besides the headline. Each holds 2^27 complex
    elements of f32 (1 GiB; f64 and the 16-bit / Q15 types: the same bytes)
    unless its configuration fixes the batch: the reference's CPU-runnable
    case (1024 x 4096), the prime and mixed-radix lengths (B = floor(2^26/N)),
    and the differentiable-transform layout (N=4096, B=8192, input stride 2,
    distance 2N+64)."""
    def c(kind, n, prec="f32", d=-1, inplace=False, **kw):
        sp = {"kind": kind, "n": n, "dtype": prec, "dir": d, "inplace": inplace}
        sp.update(kw)
        return sp

    specs = [
        c("c2c", 16), c("c2c", 32), c("c2c", 16, "f16"),
        c("c2c", 256), c("c2c", 1024, batch=4096), c("c2c", 4096, d=1),
        c("c2c", 4096, inplace=True), c("c2c", 16384), c("c2c", 65536),
        c("c2c", 65536, d=1, inplace=True), c("c2c", 1 << 20),
        c("c2c", 1 << 20, d=1, inplace=True),
        c("c2c", 4096, "f64"), c("c2c", 65536, "f64"), c("c2c", 1 << 20, "f64"),
        c("c2c", 65536, "f64", d=1, inplace=True),
        c("r2c", 256), c("r2c", 4096, adjoint=True), c("c2r", 4096, d=1, adjoint=True),
        c("r2c", 65536, adjoint=True), c("c2r", 65536, d=1),
        c("r2c", 1 << 20), c("c2r", 1 << 20, d=1), c("r2c", 65536, "f64"),
        c("r2c", 65536, inplace=True),
        c("c2c", 4096, "q15"), c("c2c", 4096, "f16"),
        c("r2r", 4096, r2r="dct2", adjoint=True),
    ]
    # prime and mixed-radix lengths, B = floor(2^26 / N)
    for prec in ("f32", "f64"):
        for n in (17, 97, 1009, 65537, 15015):
            specs.append(c("c2c", n, prec, batch=(1 << 26) // n))
    # the differentiable-transform layout: input stride 2, distance 2N+64
    for kind, r2r in (("r2c", 0), ("r2r", "dct2"), ("r2r", "dst2")):
        specs.append(c(kind, 4096, batch=8192, istride=2, idist=2 * 4096 + 64, r2r=r2r,
                       adjoint=True, layout="cfg5"))
    return specs


def run_case(torch, P, sp, x, y, xr, g, args):
    """Time one case on the resident buffers; adjoint/forward time ratio when
    the spec asks for it."""
    kind, cn, prec, cdir, inplace = sp["kind"], sp["n"], sp["dtype"], sp["dir"], sp["inplace"]
    total = x.numel()
    esz = {"f64": 16, "q15": 4, "f16": 4, "bf16": 4}.get(prec, 8)  # bytes per complex
    cb = sp.get("batch") or total * 8 // esz // cn
    is_ = sp.get("istride", 1)
    in_count = cn // 2 + 1 if kind == "c2r" else cn
    out_count = cn // 2 + 1 if kind == "r2c" else cn
    id_ = sp.get("idist") or in_count * is_
    rec = {"kind": kind, "n": cn, "dtype": prec, "batch": cb,
           "dir": "fwd" if cdir < 0 else "bwd",
           "placement": "in-place" if inplace else "out-of-place"}
    if sp.get("r2r"):
        rec["r2r"] = sp["r2r"]
    if sp.get("layout"):
        rec["layout"] = f"istride={is_} idist={id_}"
    if inplace and kind in ("r2c", "c2r"):
        # FFTW-style padded real rows so that both sides fit one buffer
        id_ = 2 * (cn // 2 + 1) if kind == "r2c" else cn // 2 + 1
        od_ = cn // 2 + 1 if kind == "r2c" else 2 * (cn // 2 + 1)
    else:
        od_ = out_count
    progress(f"case {kind} n={cn} {prec} batch={cb} {rec['dir']} {rec['placement']}")
    try:
        cplan = P.Plan(kind, cn, cdir, prec, howmany=cb, istride=is_, idist=id_, odist=od_,
                       may_alias=inplace, r2r=sp.get("r2r") or 0)
    except P.PlfftError as e:
        rec["status"] = f"unsupported: {e}"
        return rec
    dst = x if inplace else y
    steps = max(10, args.steps // 2)
    cms = time_plan(torch, cplan, x, dst, steps, args.warmup)
    half = esz // 2  # bytes per real
    if kind == "c2c":
        cbytes = 2 * cn * esz * cb
        cfl = flops_c2c(cn, cb)
    elif kind == "r2r":  # n reals in, n reals out
        cbytes = 2 * cn * half * cb
        cfl = flops_c2c(cn, cb) / 2
    else:
        cbytes = (cn * half + (cn // 2 + 1) * esz) * cb
        cfl = flops_c2c(cn, cb) / 2
    gbs = cbytes / (cms * 1e-3) / 1e9
    rec.update({"ms": round(cms, 4), "GFLOP/s": round(cfl / (cms * 1e-3) / 1e9, 1),
                "GB/s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                "plan": cplan.describe()})
    if sp.get("adjoint") and not inplace:
        # adjoint / forward time: the adjoint reads the forward output's
        # layout and writes the input's
        ams = time_plan(torch, cplan, y, x, steps, args.warmup, adjoint=True)
        rec["adjoint_over_forward"] = round(ams / cms, 3)
        rec["adjoint_ms"] = round(ams, 4)
    del cplan
    if inplace or sp.get("adjoint"):  # restore the synthetic input
        xr.uniform_(-1.0, 1.0, generator=g)
        torch.cuda.synchronize()
    return rec


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--headline-only", action="store_true")
    ap.add_argument("--mixed-only", action="store_true",
                    help="run only the mixed-length batch comparison")
    ap.add_argument("--mixed-fused-only", action="store_true",
                    help="run only the fused f32 mixed batch (counter collection)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain 1 GiB device copy (counter calibration)")
    ap.add_argument("--only", default=None,
                    help="run one case kind:n:dtype[:bwd][:inplace][:r2r kind] (profiling)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): the batch "
              "would shard contiguously across GPUs, one process each, with no "
              "communication; only the one-GPU plan is built and measured here", file=sys.stderr)
        return 2

    import torch

    import plfft_amd as P

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)

    if args.mixed_only or args.mixed_fused_only:
        print(json.dumps({"mixed": bench_mixed(torch, P, max(3, args.steps // 10),
                                               min(args.warmup, 3), args.mixed_fused_only)}))
        return 0

    total = 1 << 27  # complex elements per case
    x = torch.empty(total, dtype=torch.complex64, device="cuda")
    y = torch.empty(total, dtype=torch.complex64, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    xr = x.view(torch.float32)
    xr.uniform_(-1.0, 1.0, generator=g)  # synthetic U[-1,1) real/imag parts
    torch.cuda.synchronize()

    if args.copy_probe:
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        reps = 20
        start.record()
        for _ in range(reps):
            y.copy_(x)
        end.record()
        end.synchronize()
        ms = start.elapsed_time(end) / reps
        print(json.dumps({"copy_probe_bytes_read": total * 8, "copy_probe_bytes_written": total * 8,
                          "ms": round(ms, 5), "GB/s": round(2 * total * 8 / (ms * 1e-3) / 1e9, 1)}))
        return 0

    if args.only:
        parts = args.only.split(":")
        kind, cn, prec = parts[0], int(parts[1]), parts[2]
        cdir = 1 if "bwd" in parts[3:] else -1
        inplace = "inplace" in parts[3:]
        r2r = next((q for q in parts[3:] if q in P.R2R_KINDS), 0)
        cb = total // cn if prec != "f64" else total // 2 // cn
        cplan = P.Plan(kind, cn, cdir, prec, howmany=cb, may_alias=inplace, r2r=r2r)
        cms = time_plan(torch, cplan, x, x if inplace else y, args.steps, args.warmup)
        print(json.dumps({"case": args.only, "batch": cb, "ms": round(cms, 5),
                          "plan": cplan.describe()}))
        return 0

    # ---- headline: C2C f32 forward N=4096, B=32768, out-of-place
    n, batch = 4096, total // 4096
    plan = P.Plan("c2c", n, P.PLFFT_FORWARD, "f32", howmany=batch)
    settle_ms = settle(torch, plan, x, y)
    ms = time_plan(torch, plan, x, y, args.steps, args.warmup)
    value = flops_c2c(n, batch) / (ms * 1e-3) / 1e9
    alg_bytes = 2 * n * 8 * batch
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    workload = f"c2c_f32_fwd_n{n}_b{batch}_oop"

    cases = []
    if not args.headline_only:
        for spec in case_specs(total):
            cases.append(run_case(torch, P, spec, x, y, xr, g, args))

    out = {
        "metric": "batched C2C f32 GFLOP/s (5 N log2 N) and %HBM peak",
        "value": round(value, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device U[-1,1) real and imaginary parts, torch generator seed 0)",
        "settle_ms": round(settle_ms, 1),
        "timing": "untimed settle (the GPU leaves its idle power state), then the warmup "
                  "steps, then exactly `steps` launches between HIP events on the launch "
                  "stream; ms_per_step is their mean",
        "config": {"workload": workload, "n": n, "batch": batch, "direction": "forward",
                   "placement": "out-of-place", "layout": "contiguous, stride 1, distance N",
                   "parallelism": "single GPU (batch-sharded design, no collectives)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(workload)},
        "plan": plan.describe(),
        "cases": cases,
    }
    if not args.headline_only and not args.no_mixed:
        del x, y, xr
        torch.cuda.empty_cache()
        out["mixed"] = bench_mixed(torch, P, max(3, args.steps // 10), min(args.warmup, 3))
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(n)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": str(e)}
        try:
            out["cpu_scipy"] = cpu_scipy(n)
        except Exception as e:
            out["cpu_scipy"] = {"value": None, "error": str(e)}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
