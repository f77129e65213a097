#!/usr/bin/env python3
"""Benchmark of the transform hot path on one MI355X.

Workload (the 2D configuration the headline metric is quoted on):
DoubleFFT_2D(4096, 4096) -- complexForward on a 4096x4096 FP64 interleaved
complex array plus realForward on a 4096x4096 FP64 real array, both in place,
entries i.i.d. uniform[-1, 1).  One step = those two calls.  Every step works
on its own freshly generated input copies (resident in HBM before the timed
region), so no step transforms the output of an earlier one.

The same JSON line carries the metric's second half under "split_fp16": the
batch of 16384 x 4096-point FP32 transforms (standard normal) as plain FP32
Stockham and as split-FP16 MFMA radix-16 stages, each timed the same way,
with its rel-l2 error against numpy's FP64 FFT on 64 sampled lines and the
MFMA utilisation measured by rocprofv3 counters.

Prints one JSON line: GFLOP/s with the 5 N log2 N (complex) / 2.5 N log2 N
(real input) flop convention, the roofline of the dominant kernel (its
per-launch duration measured live with HIP events on its launch stream), and
a CPU baseline: the oracle's C port of the transforms timed on this host's
cores over a bounded sample.

Measurement helpers (no GPU needed for the last one):
  --calib-copy        three 256 MiB device copies first (the calibration
                      dispatches of a rocprofv3 --pmc run)
  --launch-log PATH   the library appends every launch label to PATH in
                      launch order (JT_LAUNCH_LOG)
  --summarize-pmc DIR turns rocprofv3 FETCH_SIZE / WRITE_SIZE runs saved under
                      DIR into per-kernel HBM bytes per launch, merged into
                      profiles/pmc_traffic.json
"""
import argparse
import csv
import glob
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
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "pmc_traffic.json")
MFMA_JSON = os.path.join(ROOT, "profiles", "mfma_util.json")
CALIB_BYTES = 256 << 20


def flops_per_step():
    n = ROWS * COLS
    return 5.0 * n * math.log2(n) + 2.5 * n * math.log2(n)


def family_key(label):
    """A launch label without its per-launch sizes: the kernel instantiation
    rocprofv3 reports as one kernel (a ragged batch: its precision only, the
    tuned schedule may differ between runs)."""
    m = re.match(r"ragged<(f64|f32),", label)
    if m:
        return f"ragged<{m.group(1)}>"
    return re.sub(r",(lines|es|n1|n2)=\d+", "", label)


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


def ragged_alg_bytes(prec, lens):
    """Algorithmic HBM bytes of one ragged-batch launch: every point read and
    written once per pass, one pass up to the single-pass limit (4096 points
    FP64, 8192 FP32), two (four-step) above; Bluestein lengths counted at n."""
    import numpy as np
    es = 16 if prec == "f64" else 8
    single = 4096 if prec == "f64" else 8192
    passes = np.where(lens <= single, 1, 2)
    return float(np.sum(2.0 * es * lens * passes))


def load_json(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def load_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 PMC summary."""
    ent = load_json(TRAFFIC_JSON).get("kernels", {}).get(kernel)
    return ent.get("hbm_bytes_per_launch") if ent else None


def cpu_baseline(budget_s):
    """The oracle's C port (OpenMP over lines) on the same workload shape."""
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
    return {
        "value": flops_per_step() / best / 1e9,
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"full step (complexForward + realForward 4096x4096 FP64), {runs} runs, "
                   f"minimum time {best * 1e3:.1f} ms, OpenMP threads={threads}; JTransforms "
                   f"itself needs a JVM, absent on this host"),
    }


def roofline_from_profile(prof, alg_override=None):
    """Dominant kernel instantiation (largest total time) and its achieved
    HBM rate: algorithmic bytes per launch / average launch duration."""
    fam = {}
    for k, (c, ms) in prof.items():
        f = fam.setdefault(family_key(k), [0, 0.0, 0.0])
        f[0] += c
        f[1] += ms
        f[2] += (alg_bytes(k) or 0.0) * c
    name, (launches, ms, byt) = max(fam.items(), key=lambda kv: kv[1][1])
    avg_s = ms / launches * 1e-3
    per = byt / launches if byt else alg_override
    ach = per / avg_s / 1e9 if (per and avg_s > 0) else None
    kernels = {k: {"launches": c, "avg_us": m / c * 1e3,
                   "GBps": (alg_bytes(k) / (m / c * 1e-3) / 1e9) if alg_bytes(k) else None}
               for k, (c, m) in sorted(prof.items())}
    return {"bound": "hbm", "kernel": name, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS if ach else None, "traffic": load_traffic(name),
            "avg_launch_us": avg_s * 1e6, "alg_bytes_per_launch": per, "kernels": kernels}


def prewarm(torch, calls):
    """One untimed call per buffer: on a fresh box the first passes over new
    device allocations run up to 10% slower, which the few warm-up steps of a
    timed run do not cover (they touch only their own buffers)."""
    for c in calls:
        c()
    torch.cuda.synchronize()


def refill(torch, bufs, base):
    """Restore every input buffer to the same synthetic input (in place)."""
    for b in bufs:
        b.copy_(base)
    torch.cuda.synchronize()


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


def c3_forms(torch, jt, steps, warmup):
    """16384 x 4096-point complexForward, FP32, standard normal: the plain
    FP32 Stockham form and the split-FP16 MFMA form, each timed over `steps`
    fresh copies of the input; rel-l2 of each against numpy's FP64 FFT of the
    same (float) input on 64 sampled lines, first and last included."""
    import numpy as np
    n, batch = 4096, 16384
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(20110602)
    base = torch.randn(2 * n * batch, dtype=torch.float32, device=dev, generator=g)
    flops = 5.0 * n * batch * math.log2(n)
    rng = np.random.default_rng(7)
    lines = sorted(set([0, batch - 1] + rng.choice(batch, 62, replace=False).tolist()))
    xs = base.view(batch, 2 * n)[lines].cpu().numpy().astype(np.float64)
    ref = np.fft.fft(xs[:, 0::2] + 1j * xs[:, 1::2], axis=1)
    res = {}
    for form in ("plain", "split_fp16"):
        plan = jt.FFTBatch1D(n, batch, precision="f32", split_fp16=(form == "split_fp16"))
        bufs = [base.clone() for _ in range(warmup + steps)]
        prewarm(torch, [(lambda b=b: plan.complexForward(b)) for b in bufs])
        calls = [(lambda b=b: plan.complexForward(b)) for b in bufs]
        # the step time without per-launch events; kernel durations (and the
        # accuracy check) from the same steps run again with them
        refill(torch, bufs, base)
        ms, _ = timed(torch, jt, calls, steps, warmup, profile=False)
        refill(torch, bufs, base)
        ms_ev, prof = timed(torch, jt, calls, steps, warmup)
        y = bufs[-1].view(batch, 2 * n)[lines].cpu().numpy().astype(np.float64)
        got = y[:, 0::2] + 1j * y[:, 1::2]
        err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        rf = roofline_from_profile(prof)
        rf.pop("kernels")
        res[form] = {"ms_per_step": ms, "GFLOPs": flops / (ms * 1e-3) / 1e9,
                     "rel_l2_vs_fp64": err, "roofline": rf, "ms_per_step_with_kernel_events": ms_ev}
        del bufs
    mf = load_json(MFMA_JSON)
    return {"config": "16384 x 4096-point complexForward, FP32 interleaved, standard normal (torch seed "
                      "20110602); rel_l2 against numpy.fft.fft in FP64 of the same float input, 64 "
                      "sampled lines",
            "plain": res["plain"], "split_fp16": res["split_fp16"],
            "split_over_plain": res["plain"]["ms_per_step"] / res["split_fp16"]["ms_per_step"],
            "mfma_util": mf.get("mfma_util"), "mfma_util_source": mf.get("source")}


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

    if args.workload == "batch4096":
        out.update(c3_forms(torch, jt, K, W))
        out["dtype"] = "f32"
        return out

    if args.workload == "fft3d":
        dims = (512, 512, 256)
        n_pts = 512 * 512 * 256
        flops = 5.0 * n_pts * math.log2(n_pts)
        base = torch.rand(2 * n_pts, dtype=torch.float32, device=dev, generator=g) * 2 - 1
        res = {}
        ref = base.double()
        jt.DoubleFFT_3D(*dims).complexForward(ref)
        for form in ("plain", "split_fp16"):
            plan = jt.FloatFFT_3D(*dims)
            if form == "split_fp16":
                plan.set_split_fp16(True)
            bufs = [base.clone() for _ in range(W + K)]
            calls = [(lambda b=b: plan.complexForward(b)) for b in bufs]
            prewarm(torch, calls)
            refill(torch, bufs, base)
            ms, prof = timed(torch, jt, calls, K, W)
            res[form] = {"ms_per_step": ms, "GFLOPs": flops / (ms * 1e-3) / 1e9,
                         "rel_l2_vs_fp64_gpu": rel(bufs[-1], ref), "roofline": roofline_from_profile(prof)}
            del bufs
        out.update({"config": "FloatFFT_3D 512x512x256 complexForward, uniform[-1,1)", "dtype": "f32",
                    "plain": res["plain"], "split_fp16": res["split_fp16"],
                    "split_over_plain": res["plain"]["ms_per_step"] / res["split_fp16"]["ms_per_step"]})
        return out

    if args.workload in ("dct3d", "dst3d"):
        dims = (512, 512, 512)
        n_pts = 512 ** 3
        plan = (jt.DoubleDCT_3D if args.workload == "dct3d" else jt.DoubleDST_3D)(*dims)
        base = torch.rand(n_pts, dtype=torch.float64, device=dev, generator=g) * 2 - 1
        bufs = [base.clone() for _ in range(W + K)]
        calls = [(lambda b=b: plan.forward(b, True)) for b in bufs]
        prewarm(torch, calls)
        refill(torch, bufs, base)
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
            byt = ragged_alg_bytes(prec, lens)
            for graph in ((False,) if args.no_graph else (False, True)):
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
                # (graph replays record no per-launch events: kernel time and
                # roofline come from the direct launches)
                rk = {k: v for k, v in prof.items() if k.startswith("ragged<")}
                kern_ms = sum(v[1] for v in rk.values()) / sum(v[0] for v in rk.values()) if rk else None
                rf = roofline_from_profile(rk, alg_override=byt) if rk else None
                if rf:
                    rf.pop("kernels")
                res[key] = {"ms_per_step": ms, "GFLOPs": flops / (ms * 1e-3) / 1e9,
                            "step_GBps_alg": byt / (ms * 1e-3) / 1e9, "plan_s": t_plan,
                            "schedule": plan.describe(), "kernel_ms": kern_ms, "roofline": rf}
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


# ---------------------------------------------------------------------------
# rocprofv3 PMC summaries
# ---------------------------------------------------------------------------
def _pmc_dispatches(run_dir, counter):
    """[(kernel name, counter value)] per dispatch, in dispatch order, from
    the counter_collection CSV(s) of one rocprofv3 --pmc run."""
    per = {}
    for path in glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                d = int(row["Dispatch_Id"])
                ent = per.setdefault(d, [row["Kernel_Name"], 0.0])
                ent[1] += float(row["Counter_Value"])
    return [tuple(per[d]) for d in sorted(per)]


def _pmc_run(run_dir, log_path, counter):
    """Scale factor from the calibration copies and per-family raw counter
    sums for one run; library dispatches (kernel names in namespace jt) are
    matched in order with the launch log."""
    disp = _pmc_dispatches(run_dir, counter)
    copies = [v for k, v in disp if "copyBuffer" in k][:3]
    if len(copies) < 3:
        raise SystemExit(f"{run_dir}: calibration copies not found (run bench.py with --calib-copy)")
    scale = CALIB_BYTES / (sorted(copies)[1] * 1024.0)  # KiB per dispatch -> bytes
    labels = [l.strip() for l in open(log_path) if l.strip()]
    ours = [(k, v) for k, v in disp if "jt::" in k]
    if len(ours) != len(labels):
        raise SystemExit(f"{run_dir}: {len(ours)} library dispatches but {len(labels)} logged launches")
    fam = {}
    for (k, v), lab in zip(ours, labels):
        f = fam.setdefault(family_key(lab), [])
        f.append((v * 1024.0 * scale, alg_bytes(lab), lab))
    return scale, fam


def _benched_launches(launches):
    """The launches of the schedule that was timed.  A ragged batch's label
    carries its schedule, and its family also holds the autotuner's trial
    launches: keep those of the label launched most often (the timed steps
    repeat the tuned schedule).  Other families are returned whole."""
    counts = {}
    for _, _, lab in launches:
        counts[lab] = counts.get(lab, 0) + 1
    if not launches or not launches[0][2].startswith("ragged<"):
        return launches
    best = max(counts, key=counts.get)
    return [x for x in launches if x[2] == best]


MIN_LISTED_BYTES = 32 << 20


def summarize_pmc(pmc_dir):
    """For every workload W with runs pmc_dir/W_fetch (FETCH_SIZE) and
    pmc_dir/W_write (WRITE_SIZE) and launch logs W.fetch.log / W.write.log:
    HBM bytes per launch of each kernel family with at least 32 MiB of
    traffic per launch, merged into profiles/pmc_traffic.json; prints a
    markdown table per workload."""
    data = load_json(TRAFFIC_JSON) or {}
    data.setdefault("kernels", {})
    for fetch_dir in sorted(glob.glob(os.path.join(pmc_dir, "*_fetch"))):
        w = os.path.basename(fetch_dir)[:-len("_fetch")]
        write_dir = os.path.join(pmc_dir, f"{w}_write")
        fs, ff = _pmc_run(fetch_dir, os.path.join(pmc_dir, f"{w}.fetch.log"), "FETCH_SIZE")
        ws, wf = _pmc_run(write_dir, os.path.join(pmc_dir, f"{w}.write.log"), "WRITE_SIZE")
        print(f"\n### {w}\n\nCalibration: a 256 MiB device copy reads FETCH_SIZE x{1 / fs:.3f} and "
              f"WRITE_SIZE x{1 / ws:.3f} of its size (scales x{fs:.3f}, x{ws:.3f}).\n")
        print("| kernel family | launches | fetch (MiB/launch) | write (MiB/launch) | HBM traffic (MiB) "
              "| algorithmic (MiB) | traffic / alg |")
        print("|---|---|---|---|---|---|---|")
        for key in sorted(set(ff) & set(wf)):
            fl, wl = _benched_launches(ff[key]), _benched_launches(wf[key])
            fe = sum(v for v, _, _ in fl) / len(fl)
            wr = sum(v for v, _, _ in wl) / len(wl)
            al = [a for _, a, _ in fl if a]
            alg = sum(al) / len(al) if al else None
            m = re.match(r"ragged<(f64|f32)>", key)
            if m and w == "ragged":
                from jtransforms_amd.workloads import ragged_lengths
                alg = ragged_alg_bytes(m.group(1), ragged_lengths(20000))
            tot = fe + wr
            if tot < MIN_LISTED_BYTES:
                continue  # plan-creation kernels (Bluestein spectra, chirps) and tiny helpers
            data["kernels"][key] = {"hbm_bytes_per_launch": tot, "fetch_bytes": fe, "write_bytes": wr,
                                    "alg_bytes_per_launch": alg, "launches": len(fl),
                                    "traffic_over_alg": (tot / alg) if alg else None, "workload": w}
            if len(fl) != len(ff[key]):
                data["kernels"][key]["schedule"] = fl[0][2]
            print(f"| `{key}` | {len(fl)} | {fe / 2**20:.1f} | {wr / 2**20:.1f} | {tot / 2**20:.1f} | "
                  f"{(alg / 2**20) if alg else float('nan'):.1f} | {(tot / alg) if alg else float('nan'):.3f} |")
    data["calibration"] = {"note": "per run: scale = 256 MiB / counter reading of a 256 MiB device copy "
                                   "(__amd_rocclr_copyBuffer, bench.py --calib-copy) in the same run"}
    with open(TRAFFIC_JSON, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-split", action="store_true", help="leave out the split_fp16 sub-object")
    ap.add_argument("--workload", default="fft2d",
                    choices=["fft2d", "batch4096", "fft3d", "dct3d", "dst3d", "fft1d", "ragged"],
                    help="fft2d (default): the headline 4096x4096 FP64 line; the others "
                         "measure the remaining configurations")
    ap.add_argument("--no-graph", action="store_true",
                    help="ragged: direct launches only (a profiler then sees one labelled launch per dispatch)")
    ap.add_argument("--calib-copy", action="store_true")
    ap.add_argument("--launch-log", default=None)
    ap.add_argument("--summarize-pmc", default=None, metavar="DIR")
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the transforms "
              "here are single-device and no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)
    if args.launch_log:
        os.environ["JT_LAUNCH_LOG"] = os.path.abspath(args.launch_log)

    import torch

    import jtransforms_amd as jt

    if args.calib_copy:
        src = torch.empty(CALIB_BYTES, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

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
    prewarm(torch, [(lambda b=b: (plan.complexForward(b[0]), plan.realForward(b[1]))) for b in bufs])

    def run_steps(instrumented):
        """W warm-up steps, then K timed steps on fresh input copies; with
        `instrumented` the library records HIP events around every launch."""
        refill(torch, [b[0] for b in bufs], base_c)
        refill(torch, [b[1] for b in bufs], base_r)
        for i in range(args.warmup):
            plan.complexForward(bufs[i][0])
            plan.realForward(bufs[i][1])
        torch.cuda.synchronize()
        if instrumented:
            jt.profile_begin()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.warmup, total):
            plan.complexForward(bufs[i][0])
            plan.realForward(bufs[i][1])
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        prof = jt.profile_end() if instrumented else None
        return el * 1e3 / args.steps, prof

    # the step time is measured without per-launch events; the kernel
    # durations of the roofline come from the same K steps run again with
    # the events recorded (their own step time is reported beside it)
    ms_per_step, _ = run_steps(False)
    ms_profiled, prof = run_steps(True)
    del bufs
    value = flops_per_step() / (ms_per_step * 1e-3) / 1e9

    # dominant kernel: the instantiation (launch label without its line
    # count, i.e. what rocprofv3 groups as one kernel) with the largest total
    # time; its algorithmic bytes and duration are averaged over its launches
    roofline = roofline_from_profile(prof)
    kernels = roofline.pop("kernels")
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
        "ms_per_step_with_kernel_events": ms_profiled,
    }
    if not args.no_split:
        out["split_fp16"] = c3_forms(torch, jt, args.steps, args.warmup)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
