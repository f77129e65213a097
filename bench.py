#!/usr/bin/env python3
"""Benchmark of the transform hot path on one MI355X.

Workload (the 2D configuration the headline metric is quoted on):
DoubleFFT_2D(4096, 4096) -- complexForward on a 4096x4096 FP64 interleaved
complex array plus realForward on a 4096x4096 FP64 real array, both in place,
entries i.i.d. uniform[-1, 1).  One step = those two calls.  Every step works
on its own freshly generated input copies (resident in HBM before the timed
region), so no step transforms the output of an earlier one.

Prints one JSON line: GFLOP/s with the 5 N log2 N (complex) / 2.5 N log2 N
(real input) flop convention, the roofline of the dominant kernel (its
per-launch duration measured live with HIP events on its launch stream), and
a CPU baseline: the oracle's C port of the transforms timed on this host's
cores over a bounded sample.
"""
import argparse
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("GFLOP/s (5Nlog2N) & %HBM peak: FP64 plain vs split-FP16 MFMA @ rel-ℓ2≤1e-6")
HBM_PEAK_GBS = 8000.0
ROWS = COLS = 4096


def flops_per_step():
    n = ROWS * COLS
    return 5.0 * n * math.log2(n) + 2.5 * n * math.log2(n)


def parse_kernel(name):
    """(precision bytes, N, lines) of a stockham/generic launch label."""
    m = re.match(r"(stockham|generic)<(f64|f32),N=(\d+),.*lines=(\d+),", name)
    if not m:
        return None
    return (8 if m.group(2) == "f64" else 4), int(m.group(3)), int(m.group(4))


def alg_bytes(name):
    """Algorithmic HBM bytes of one launch: every point read once and written
    once (N complex points per line, or 2N reals for the real-data tags)."""
    p = parse_kernel(name)
    if p is None:
        return None
    es, n, lines = p
    return 2.0 * lines * n * 2 * es


def load_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    ent = data.get("kernels", {}).get(kernel)
    if not ent:
        return None
    return ent.get("hbm_bytes_per_launch")


def cpu_baseline(budget_s):
    """The oracle's C port (OpenMP over lines) on the same workload shape."""
    import numpy as np
    from oracle import oracle as orc
    threads = orc.num_threads()
    x = orc.jrandom_doubles(20110602, 2 * ROWS * COLS)
    y = orc.jrandom_doubles(20110603, ROWS * COLS)
    best = float("inf")
    runs = 0
    t_begin = time.perf_counter()
    while runs < 2 or (time.perf_counter() - t_begin < budget_s and runs < 50):
        t0 = time.perf_counter()
        orc.fftnd_complex([ROWS, COLS], x, -1)
        orc.fft2d_real_forward(ROWS, COLS, y)
        dt = time.perf_counter() - t0
        best = min(best, dt)
        runs += 1
    del np
    return {
        "value": flops_per_step() / best / 1e9,
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"full step (complexForward + realForward 4096x4096 FP64), {runs} runs, "
                   f"minimum time {best * 1e3:.1f} ms, OpenMP threads={threads}; JTransforms "
                   f"itself needs a JVM, absent on this host"),
    }


def roofline_from_profile(prof):
    """Dominant kernel instantiation and its achieved HBM rate."""
    fam = {}
    for k, (c, ms) in prof.items():
        key = re.sub(r",(lines|es)=\d+", "", k)
        f = fam.setdefault(key, [0, 0.0, 0.0])
        f[0] += c
        f[1] += ms
        f[2] += (alg_bytes(k) or 0.0) * c
    name, (launches, ms, byt) = max(fam.items(), key=lambda kv: kv[1][1])
    avg_s = ms / launches * 1e-3
    per = byt / launches if byt else None
    ach = per / avg_s / 1e9 if per else None
    kernels = {k: {"launches": c, "avg_us": m / c * 1e3,
                   "GBps": (alg_bytes(k) / (m / c * 1e-3) / 1e9) if alg_bytes(k) else None}
               for k, (c, m) in sorted(prof.items())}
    return {"bound": "hbm", "kernel": name, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS if ach else None, "traffic": load_traffic(name),
            "avg_launch_us": avg_s * 1e6, "alg_bytes_per_launch": per, "kernels": kernels}


def timed(torch, jt, calls, steps, warmup, profile=True):
    """Run calls[i]() for warm-up then timed steps; returns (ms/step, profile).
    With profile=False the library records no per-launch events (None is
    returned for the profile)."""
    for i in range(warmup):
        calls[i]()
    torch.cuda.synchronize()
    if profile:
        jt.profile_begin()
    t0 = time.perf_counter()
    for i in range(warmup, warmup + steps):
        calls[i]()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return el * 1e3 / steps, (jt.profile_end() if profile else None)


def other_workload(args, torch, jt):
    """The other configurations: one JSON object each (not the headline line)."""
    import numpy as np
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(20110602)
    K, W = args.steps, args.warmup
    out = {"workload": args.workload, "steps": K, "warmup": W}

    def rel(a, b):
        return float(torch.linalg.vector_norm((a.double() - b.double())) / torch.linalg.vector_norm(b.double()))

    if args.workload in ("batch4096", "fft3d"):
        if args.workload == "batch4096":
            n_pts, dims = 16384 * 4096, None
            mk = lambda: jt.FFTBatch1D(4096, 16384, precision="f32")
            mk64 = lambda: jt.FFTBatch1D(4096, 16384, precision="f64")
            flops = 5.0 * n_pts * math.log2(4096)
            desc = "16384 x 4096-point complexForward, FP32, standard normal"
            base = torch.randn(2 * n_pts, dtype=torch.float32, device=dev, generator=g)
        else:
            dims = (512, 512, 256)
            n_pts = 512 * 512 * 256
            mk = lambda: jt.FloatFFT_3D(*dims)
            mk64 = lambda: jt.DoubleFFT_3D(*dims)
            flops = 5.0 * n_pts * math.log2(n_pts)
            desc = "FloatFFT_3D 512x512x256 complexForward, uniform[-1,1)"
            base = torch.rand(2 * n_pts, dtype=torch.float32, device=dev, generator=g) * 2 - 1
        res = {}
        ref = base.double()
        mk64().complexForward(ref)
        for form in ("plain", "split_fp16"):
            plan = mk()
            if form == "split_fp16":
                plan.set_split_fp16(True)
            bufs = [base.clone() for _ in range(W + K)]
            calls = [(lambda b=b: plan.complexForward(b)) for b in bufs]
            ms, prof = timed(torch, jt, calls, K, W)
            chk = base.clone()
            plan.complexForward(chk)
            torch.cuda.synchronize()
            res[form] = {"ms_per_step": ms, "GFLOPs": flops / (ms * 1e-3) / 1e9,
                         "rel_l2_vs_fp64_gpu": rel(chk, ref), "roofline": roofline_from_profile(prof)}
            del bufs
        out.update({"config": desc, "dtype": "f32", "plain": res["plain"], "split_fp16": res["split_fp16"],
                    "split_over_plain": res["plain"]["ms_per_step"] / res["split_fp16"]["ms_per_step"]})
        return out

    if args.workload in ("dct3d", "dst3d"):
        dims = (512, 512, 512)
        n_pts = 512 ** 3
        plan = (jt.DoubleDCT_3D if args.workload == "dct3d" else jt.DoubleDST_3D)(*dims)
        base = torch.rand(n_pts, dtype=torch.float64, device=dev, generator=g) * 2 - 1
        bufs = [base.clone() for _ in range(W + K)]
        calls = [(lambda b=b: plan.forward(b, True)) for b in bufs]
        ms, prof = timed(torch, jt, calls, K, W)
        chk = base.clone()
        plan.forward(chk, True)
        plan.inverse(chk, True)
        torch.cuda.synchronize()
        flops = 2.5 * n_pts * math.log2(n_pts)
        out.update({"config": f"{plan.__class__.__name__} 512^3 forward(scale=true), FP64", "dtype": "f64",
                    "ms_per_step": ms, "GFLOPs": flops / (ms * 1e-3) / 1e9,
                    "round_trip_rel_l2": rel(chk, base), "roofline": roofline_from_profile(prof),
                    "step_GBps": 3 * 2 * n_pts * 8 / (ms * 1e-3) / 1e9})
        return out

    if args.workload == "ragged":
        from jtransforms_amd.workloads import ragged_lengths
        lens = ragged_lengths(20000)
        flops = float(np.sum(5.0 * lens * np.log2(lens)))
        res = {}
        for prec in ("f64", "f32"):
            dt = torch.float64 if prec == "f64" else torch.float32
            es = 16 if prec == "f64" else 8
            # algorithmic bytes: read + write once per pass; 2 passes above 4096 points
            passes = np.where(lens <= 4096, 1, 2)
            byt = float(np.sum(2.0 * es * lens * passes))
            for graph in (False, True):
                print(f"ragged {prec} graph={graph}", file=sys.stderr, flush=True)
                t0 = time.perf_counter()
                # the first creation per precision runs the autotuner; the
                # second reuses the schedule it chose
                plan = jt.RaggedBatch1D(lens, precision=prec, graph=graph, autotune=True)
                t_plan = time.perf_counter() - t0
                tot = int(2 * lens.sum())
                base = torch.rand(tot, dtype=dt, device=dev, generator=g) * 2 - 1
                nb = min(W + K, 6)
                bufs = [base.clone() for _ in range(nb)]
                st = torch.cuda.Stream()
                with torch.cuda.stream(st):
                    calls = [(lambda b=bufs[i % nb]: plan.complexForward(b)) for i in range(W + K)]
                    ms, prof = timed(torch, jt, calls, K, W)
                plan.sync_status()
                key = f"{prec}_{'graph' if graph else 'direct'}"
                rk = [v for k, v in prof.items() if k.startswith("ragged<")]
                kern_ms = rk[0][1] / rk[0][0] if rk else None
                res[key] = {"ms_per_step": ms, "GFLOPs": flops / (ms * 1e-3) / 1e9,
                            "step_GBps_alg": byt / (ms * 1e-3) / 1e9, "plan_s": t_plan,
                            "schedule": plan.describe(),
                            "kernel_ms": kern_ms,
                            "kernel_GBps_alg": byt / (kern_ms * 1e-3) / 1e9 if kern_ms else None}
                del bufs, base, plan
        out.update({"config": "ragged batch of 20000 1D complex transforms, n log-uniform in [16, 65536] "
                              "(70% 2^a3^b5^c, 25% 2^k, 5% primes), one persistent launch",
                    "total_points": int(lens.sum()), "flops": flops, "results": res})
        return out

    # fft1d: n=1024 FP64, batch 1: complexForward + complexInverse(scale) + realForward
    n = 1024
    plan = jt.DoubleFFT_1D(n)
    base = torch.rand(2 * n, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    bufs = [base.clone() for _ in range(W + K)]

    def step(b):
        plan.complexForward(b)
        plan.complexInverse(b, True)
        plan.realForward(b)

    calls = [(lambda b=b: step(b)) for b in bufs]
    # latency-bound: the step is timed without the per-launch profiling
    # events, and the kernel durations come from a second, profiled pass
    ms, _ = timed(torch, jt, calls, K, W, profile=False)
    _, prof = timed(torch, jt, calls, K, W)
    out.update({"config": "DoubleFFT_1D(1024): complexForward + complexInverse(scale) + realForward, FP64",
                "dtype": "f64", "us_per_step": ms * 1e3,
                "kernel_us": {k: ms_ / c * 1e3 for k, (c, ms_) in prof.items()}})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="fft2d",
                    choices=["fft2d", "batch4096", "fft3d", "dct3d", "dst3d", "fft1d", "ragged"],
                    help="fft2d (default): the headline 4096x4096 FP64 line; the others "
                         "measure the remaining configurations")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the transforms "
              "here are single-device and no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)

    import torch

    import jtransforms_amd as jt

    if args.workload != "fft2d":
        print(json.dumps(other_workload(args, torch, jt)))
        return

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    plan = jt.DoubleFFT_2D(ROWS, COLS)
    g = torch.Generator(device=dev)
    g.manual_seed(20110602)
    base_c = torch.rand(2 * ROWS * COLS, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    base_r = torch.rand(ROWS * COLS, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    total = args.warmup + args.steps
    bufs = [(base_c.clone(), base_r.clone()) for _ in range(total)]
    torch.cuda.synchronize()

    for i in range(args.warmup):
        plan.complexForward(bufs[i][0])
        plan.realForward(bufs[i][1])
    torch.cuda.synchronize()

    jt.profile_begin()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.warmup, total):
        plan.complexForward(bufs[i][0])
        plan.realForward(bufs[i][1])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = jt.profile_end()

    ms_per_step = elapsed * 1e3 / args.steps
    value = flops_per_step() / (ms_per_step * 1e-3) / 1e9

    # dominant kernel: the instantiation (launch label without its line
    # count, i.e. what rocprofv3 groups as one kernel) with the largest total
    # time; its algorithmic bytes and duration are averaged over its launches
    fam = {}
    for k, (c, ms) in prof.items():
        key = re.sub(r",(lines|es)=\d+", "", k)
        f = fam.setdefault(key, [0, 0.0, 0.0])
        f[0] += c
        f[1] += ms
        f[2] += (alg_bytes(k) or 0.0) * c
    dom_name, (dom_launches, dom_ms, dom_bytes) = max(fam.items(), key=lambda kv: kv[1][1])
    dom_avg_s = dom_ms / dom_launches * 1e-3
    bytes_launch = dom_bytes / dom_launches if dom_bytes else None
    achieved = bytes_launch / dom_avg_s / 1e9 if bytes_launch else None
    roofline = {
        "bound": "hbm",
        "kernel": dom_name,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
        "traffic": load_traffic(dom_name),
        "avg_launch_us": dom_avg_s * 1e6,
        "alg_bytes_per_launch": bytes_launch,
    }
    kernels = {k: {"launches": c, "avg_us": ms / c * 1e3,
                   "GBps": (alg_bytes(k) / (ms / c * 1e-3) / 1e9) if alg_bytes(k) else None}
               for k, (c, ms) in sorted(prof.items())}
    # whole-step HBM rate: each step moves 2 passes x (256 MiB + 128 MiB) x read+write
    step_bytes = 2 * 2 * (2 * ROWS * COLS * 8 + ROWS * COLS * 8)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: uniform[-1,1) FP64 from torch's generator, seed 20110602",
        "config": {
            "workload": "DoubleFFT_2D 4096x4096 FP64: complexForward + realForward, in place",
            "rows": ROWS,
            "columns": COLS,
            "flops_per_step": flops_per_step(),
            "hbm_bytes_per_step_alg": step_bytes,
            "step_GBps": step_bytes / (ms_per_step * 1e-3) / 1e9,
        },
        "roofline": roofline,
        "kernels": kernels,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
