#!/usr/bin/env python3
"""Benchmark of the mode-q tensor-times-vector product on one MI355X.

Workload (one "step"): the FP32 order-4 tensor 128^4 (2^28 elements, 1 GiB, U[-1,1] from a
seed), contracted over every mode q = 1..4 in first-order and in last-order layout -- 8 products
per step, inputs resident in HBM before timing starts.  `value` is the step's algorithmic bytes
(per product N*4 + (N/n_q)*4 + n_q*4 = 1,082,130,944 B) divided by the step time, in GB/s; the
step time comes from one event pair around K back-to-back steps.

Timing protocol (every leg): the launches of a timed region are enqueued behind a bounded spin
kernel, so the host has queued them all before the GPU reaches them, and one event pair on the
launch stream brackets the whole region; a launch's time is the region's time over its launches
(kernel time plus the sub-microsecond gap between consecutive launches).  Consecutive launches
read operands that together exceed the 256 MB last-level cache (1 GiB tensors by themselves;
smaller ones rotate over copies), so every launch streams A from HBM.

Also reported on the same line:
  roofline      -- the slowest of the 8 products: algorithmic bytes per launch / its mean launch
                   time (K launches back to back, best of three such regions) against the 8 TB/s
                   HBM peak; traffic from the committed rocprofv3 counters
  per_case      -- GB/s and fraction of peak for each (q, layout), timed the same way
  read_ceiling  -- a pure 16-byte read stream over the same 1 GiB; mixed_ceilings -- the same read
                   stream plus a write stream of 1/128, 1/16 or 1/4 of it (the output fractions of
                   128^4, 16^7 and the order-8 q = 4 product); each the faster of a persistent
                   chunked stream and a compact sweep of short-lived blocks
  precision     -- BF16 and FP8-e4m3 storage of A: speedup over the FP32 kernel per case (both timed
                   with the per-case protocol) and the relative L2 error of C against the FP64 oracle on
                   the original data (and on the dequantised data)
  cpu_baseline  -- the reference's CPU path (restated in oracle/: OpenMP over slices + OpenBLAS
                   gemv, par_loop/subtensor/all) on the host's physical cores, on the same inputs

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-precision]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

PEAK_BPS = 8.0e12
SHAPE = [128, 128, 128, 128]
N = 1 << 28


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def case_bytes(nq, elem=4, c_elem=4, extra=0):
    return N * elem + (N // nq) * c_elem + nq * 4 + extra


def spin(torch, ttv_amd, us, sink):
    """Holds the GPU stream for about `us` microseconds, so that the launches enqueued after it are
    all queued before the GPU reaches them: no launch then waits for the host, and event pairs
    time the kernels alone."""
    import ctypes
    iters = int(min(100000, max(1, us / 3.4)))
    ttv_amd.check(ttv_amd.lib().ttvx_bench_spin(iters, ctypes.c_void_p(sink.data_ptr()),
                                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))


def timed_cycle(torch, ttv_amd, runs, reps, sink, us_per_launch=200.0):
    """Mean time of one launch in ms: `reps` launches back to back on the current stream, cycling
    through `runs` (callables that each enqueue one launch, typically on their own copy of the
    operands), behind a spin, with one event pair around all of them."""
    for r in runs:
        r()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    spin(torch, ttv_amd, reps * us_per_launch, sink)
    e0.record(stream)
    for i in range(reps):
        runs[i % len(runs)]()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def copies_for(t, min_bytes=1 << 30):
    """t and enough clones of it that cycling through them reads at least min_bytes (> the
    last-level cache) between two reads of the same copy."""
    n = max(1, -(-min_bytes // (t.numel() * t.element_size())))
    return [t] + [t.clone() for _ in range(n - 1)]


def traffic_from_profiles(name="traffic_ttv_f32_128x4.json"):
    """HBM bytes per launch of the profiled kernels measured with rocprofv3 counters (committed
    summary), or None."""
    path = os.path.join(REPO, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


CASE_ORDER = [(lname, q) for lname in ("first", "last") for q in range(1, 5)]
FP8_FIBERS_PER_SCALE = 32                        # ttv_amd.default_fibers_per_scale(128)
FP8_SCALES = (N // 128) // FP8_FIBERS_PER_SCALE  # one float per 4096 elements


ORDER8_SHAPE = [4096, 4, 8, 4, 8, 4, 2, 64]
ORDER8_N = 1 << 31


def profile_cases(kind):
    """(label, algorithmic bytes per launch) of the products `bench.py --profile KIND` dispatches,
    in dispatch order (each 3 times)."""
    if kind == "f32":
        return [("%s q=%d" % (lname, q), case_bytes(128)) for lname, q in CASE_ORDER]
    if kind == "other":
        out = [("16^7 %s q=%d" % (lname, q), N * 4 + (N // 16) * 4 + 16 * 4)
               for lname in ("first", "last") for q in range(1, 8)]
        n3 = 1 << 24
        out.append(("fp64 256^3 q=2", n3 * 8 + 65536 * 8 + 256 * 8))
        for q in (1, 4, 8):
            nq = ORDER8_SHAPE[q - 1]
            out.append(("order-8 q=%d" % q, ORDER8_N * 4 + (ORDER8_N // nq) * 4 + nq * 4))
        return out
    out = []
    for lname, q in CASE_ORDER:
        out.append(("bf16 %s q=%d" % (lname, q), case_bytes(128, elem=2)))
        out.append(("fp8 %s q=%d" % (lname, q), case_bytes(128, elem=1, extra=FP8_SCALES * 4)))
    return out


def other_profile_runs(torch, ttv_amd, Aflat):
    """(label, launch) of the `--profile other` products, operands and outputs allocated up front,
    in profile_cases("other") order: FP32 16^7 (every q, first/last-order, on the 128^4 data), FP64
    256^3 q = 2, FP32 order-8 (the benchmark's seeded layout) q = 1, 4, 8."""
    import numpy as np
    from ttv_amd.meta import generate_strides
    dev = Aflat.device
    runs = []
    sh = [16] * 7
    for lname, pia in (("first", list(range(1, 8))), ("last", list(range(7, 0, -1)))):
        A = Aflat.as_strided(sh, generate_strides(sh, pia))
        for q in range(1, 8):
            b = torch.rand(16, device=dev)
            C = ttv_amd.ttv(q, A, b)
            runs.append(("16^7 %s q=%d" % (lname, q), lambda q=q, A=A, b=b, C=C: ttv_amd.ttv(q, A, b, out=C)))
    n3 = 1 << 24
    A64 = Aflat[:n3].double().view(256, 256, 256).permute(2, 1, 0)
    b64 = torch.rand(256, device=dev, dtype=torch.float64)
    C64 = ttv_amd.ttv(2, A64, b64)
    runs.append(("fp64 256^3 q=2", lambda: ttv_amd.ttv(2, A64, b64, out=C64)))
    pia8 = [int(x) for x in np.random.default_rng(31).permutation(8) + 1]
    A8 = torch.empty(ORDER8_N, dtype=torch.float32, device=dev).uniform_(-1, 1).as_strided(
        ORDER8_SHAPE, generate_strides(ORDER8_SHAPE, pia8))
    for q in (1, 4, 8):
        b = torch.rand(ORDER8_SHAPE[q - 1], device=dev)
        C = ttv_amd.ttv(q, A8, b)
        runs.append(("order-8 q=%d" % q, lambda q=q, b=b, C=C: ttv_amd.ttv(q, A8, b, out=C)))
    torch.cuda.synchronize()
    return runs


def summarize_profiles(d, traffic_out=None, kind="f32"):
    """Summary text of the three rocprofv3 runs of `bench.py --profile KIND` under directory d
    (prof_trace: --kernel-trace --stats; prof_fetch: --pmc FETCH_SIZE; prof_write: --pmc WRITE_SIZE)
    and, with traffic_out, the per-case HBM traffic JSON.  Counters are calibrated with the plain
    copy kernel of the same sequence (1 GiB read, 1 GiB written per launch)."""
    import csv
    import re

    def short(n):
        return re.sub(r"\(.*", "", n).replace("void ", "")

    def rows(path):
        with open(path) as f:
            return list(csv.DictReader(f))

    def is_product(name):
        return any(k in name for k in ("k_rows", "k_dot", "k_slab", "k_col", "k_strided"))

    out = []
    stats = rows(os.path.join(d, "prof_trace", "run_kernel_stats.csv"))
    out.append("Kernel statistics (run_kernel_stats.csv):")
    for r in stats:
        if "ttvx" in r["Name"]:
            out.append("  %-58s calls=%3s avg_us=%8.1f min_us=%8.1f max_us=%8.1f" % (
                short(r["Name"]), r["Calls"], float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
                float(r["MaxNs"]) / 1e3))

    def after_streams(rs):
        """ttvx dispatches in dispatch order; the products timed are those after the stream kernels
        (the ones before them only allocated the outputs)"""
        rs = sorted((r for r in rs if "ttvx" in r["Kernel_Name"]), key=lambda r: int(r["Dispatch_Id"]))
        last = max(i for i, r in enumerate(rs) if "stream" in r["Kernel_Name"])
        return rs, [r for r in rs[last + 1:] if is_product(r["Kernel_Name"])]

    _, prod_t = after_streams(rows(os.path.join(d, "prof_trace", "run_kernel_trace.csv")))

    def counters(name):
        rs, prod = after_streams(rows(os.path.join(d, "prof_" + name, "run_counter_collection.csv")))
        copy = [float(r["Counter_Value"]) for r in rs if "k_copy_stream" in r["Kernel_Name"]]
        return sum(copy) / len(copy), [(short(r["Kernel_Name"]), float(r["Counter_Value"])) for r in prod]

    fetch_copy, fetch = counters("fetch")
    write_copy, write = counters("write")
    gib_kib = float(1 << 20)
    fscale, wscale = gib_kib / fetch_copy, gib_kib / write_copy
    out.append("")
    out.append("Counter calibration with the plain copy of 1 GiB (reads 1,048,576 KiB, writes 1,048,576 KiB):")
    out.append("  FETCH_SIZE read %.1f KiB -> scale %.4f ; WRITE_SIZE read %.1f KiB -> scale %.4f" % (
        fetch_copy, fscale, write_copy, wscale))
    plan = profile_cases(kind)
    reps = len(prod_t) // len(plan)
    assert reps >= 1 and len(prod_t) == reps * len(plan) and len(fetch) == len(prod_t) and len(write) == len(prod_t), \
        "expected %d product dispatches per run" % (reps * len(plan))
    cases = []
    out.append("")
    out.append("Per case (mean of %d launches): trace duration, HBM traffic from the scaled counters, vs the "
               "algorithmic bytes:" % reps)
    for i, (label, alg) in enumerate(plan):
        sl = slice(i * reps, (i + 1) * reps)
        us = sum((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in prod_t[sl]) / reps
        fb = sum(v for _, v in fetch[sl]) / reps * 1024 * fscale
        wb = sum(v for _, v in write[sl]) / reps * 1024 * wscale
        name = short(prod_t[i * reps]["Kernel_Name"])
        entry = dict(label=label, kernel=name, avg_us=round(us, 2), algorithmic_bytes=alg, fetch_bytes=round(fb),
                     write_bytes=round(wb), hbm_bytes_per_launch=round(fb + wb),
                     traffic_over_algorithmic=round((fb + wb) / alg, 4), gbs=round(alg / (us * 1e-6) / 1e9, 1))
        if kind == "f32":
            entry["layout"], entry["q"] = CASE_ORDER[i]
        cases.append(entry)
        out.append("  %-15s %-38s %7.1f us  fetch %.4e B  write %.4e B  (%.4f x algorithmic %d B, %.1f GB/s)" % (
            label, name.replace("ttvx::", ""), us, fb, wb, (fb + wb) / alg, alg, alg / (us * 1e-6) / 1e9))
    if traffic_out:
        workload = {"f32": "FP32 128^4, q=1..4 x {first,last}-order",
                    "precision": "128^4 with A in BF16 and FP8-e4m3 (one scale per 4096 elements), q=1..4 x "
                                 "{first,last}-order",
                    "other": "FP32 16^7 (q=1..7 x {first,last}-order), FP64 256^3 q=2, FP32 order-8 "
                             "(4096,4,8,4,8,4,2,64) q=1,4,8"}[kind]
        with open(traffic_out, "w") as f:
            json.dump(dict(workload=workload, fetch_scale=round(fscale, 4), write_scale=round(wscale, 4), cases=cases),
                      f, indent=1)
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the 16^7 FP32 and 256^3 FP64 cases")
    ap.add_argument("--no-order8", action="store_true", help="skip the order-8 precision configuration")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--profile", choices=("f32", "precision", "other"),
                    help="short fixed dispatch sequence for rocprofv3: copy and read streams, then each of the "
                         "8 FP32 products (f32), the 8 BF16 and 8 FP8 products (precision), or the FP32 16^7, "
                         "FP64 256^3 and FP32 order-8 products (other) 3 times; prints no JSON")
    ap.add_argument("--summarize", metavar="DIR",
                    help="print the summary of the rocprofv3 runs of --profile under DIR (no GPU needed)")
    ap.add_argument("--kind", choices=("f32", "precision", "other"), default="f32",
                    help="with --summarize: the --profile kind")
    ap.add_argument("--traffic-out", metavar="JSON", help="with --summarize: write the per-case traffic here")
    args = ap.parse_args()
    if args.summarize:
        print(summarize_profiles(args.summarize, args.traffic_out, args.kind))
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are out of scope (the mode-q product "
              "shards over free modes with no exchange)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import ttv_amd
    from ttv_amd.meta import generate_strides

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    gen = torch.Generator(device=dev)
    gen.manual_seed(20250101)
    log("generating A (2^28 FP32, U[-1,1]) on the GPU")
    Aflat = torch.rand(N, generator=gen, device=dev, dtype=torch.float32).mul_(2).sub_(1)
    layouts = {"first": [1, 2, 3, 4], "last": [4, 3, 2, 1]}
    cases = []
    for lname, pia in layouts.items():
        wa = generate_strides(SHAPE, pia)
        A = Aflat.as_strided(SHAPE, wa)
        for q in range(1, 5):
            b = torch.rand(128, generator=gen, device=dev, dtype=torch.float32).mul_(2).sub_(1)
            info = ttv_amd.plan_info(q, SHAPE, wa, pia)
            C = ttv_amd.ttv(q, A, b)             # allocates C once; reused in the timed loop
            cases.append(dict(layout=lname, pia=pia, wa=wa, q=q, A=A, b=b, C=C, variant=info["variant"],
                              kernel=ttv_amd.variant_names()[info["variant"]], bytes=case_bytes(128)))
    torch.cuda.synchronize()

    def run_case(c):
        ttv_amd.ttv(c["q"], c["A"], c["b"], out=c["C"])

    if args.profile:
        L = ttv_amd.lib()
        # operands and outputs first: only the stream kernels and the profiled products follow
        prep = []
        if args.profile == "precision":
            Abf = ttv_amd.quantize_bf16(Aflat)
            for c in cases:
                q, wa = c["q"], c["wa"]
                codes, scales, fps = ttv_amd.quantize_fp8(q, Aflat.as_strided(SHAPE, wa))
                assert fps == FP8_FIBERS_PER_SCALE and scales.numel() == FP8_SCALES
                Aq = Abf.as_strided(SHAPE, wa)
                prep.append((c, Aq, codes, scales, fps, ttv_amd.ttv_bf16(q, Aq, c["b"]),
                             ttv_amd.ttv_fp8(q, codes, scales, fps, c["b"])))
        other = other_profile_runs(torch, ttv_amd, Aflat) if args.profile == "other" else []
        dst = torch.empty_like(Aflat)
        partial = torch.empty(256 * 32, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        for _ in range(3):
            ttv_amd.check(L.ttvx_bench_copy(Aflat.data_ptr(), dst.data_ptr(), N * 4, stream.cuda_stream))
        for _ in range(3):
            ttv_amd.check(L.ttvx_bench_read(Aflat.data_ptr(), N * 4, partial.data_ptr(), partial.numel(),
                                            stream.cuda_stream))
        if args.profile == "f32":
            for c in cases:
                log("profile case", c["layout"], c["q"], c["kernel"])
                for _ in range(3):
                    run_case(c)
        elif args.profile == "other":
            for label, run in other:
                log("profile case", label)
                for _ in range(3):
                    run()
        else:
            for c, Aq, codes, scales, fps, Cb, C8 in prep:
                log("profile case bf16, fp8", c["layout"], c["q"])
                for _ in range(3):
                    ttv_amd.ttv_bf16(c["q"], Aq, c["b"], out=Cb)
                for _ in range(3):
                    ttv_amd.ttv_fp8(c["q"], codes, scales, fps, c["b"], out=C8)
        torch.cuda.synchronize()
        log("profile sequence done")
        return

    for _ in range(args.warmup):
        for c in cases:
            run_case(c)
    torch.cuda.synchronize()

    # timed region: K steps back to back on the launch stream, one event pair around all of them
    # (an event record between launches is a packet of its own and costs ~3 us per record)
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    sink = torch.zeros(1, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    spin(torch, ttv_amd, args.steps * len(cases) * 200.0, sink)     # the host enqueues the whole region
    t_start.record(stream)
    for s in range(args.steps):
        for c in cases:
            run_case(c)
    t_end.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_end)
    step_bytes = sum(c["bytes"] for c in cases)
    value = step_bytes * args.steps / (total_ms * 1e-3) / 1e9

    # per-case regions: each product K times back to back between one event pair (its mean launch
    # time); all 8 regions are enqueued behind one spin; the best of three such passes (a region can
    # catch a stall of the box)
    for c in cases:
        c["ms"] = float("inf")
    for _ in range(3):
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in cases]
        spin(torch, ttv_amd, args.steps * len(cases) * 200.0, sink)
        for c, (e0, e1) in zip(cases, evs):
            e0.record(stream)
            for s in range(args.steps):
                run_case(c)
            e1.record(stream)
        torch.cuda.synchronize()
        for c, (e0, e1) in zip(cases, evs):
            c["ms"] = min(c["ms"], e0.elapsed_time(e1) / args.steps)
    for c in cases:
        c["gbs"] = c["bytes"] / (c["ms"] * 1e-3) / 1e9
    per_case = [dict(layout=c["layout"], q=c["q"], kernel=c["kernel"], ms=round(c["ms"], 4),
                     gbs=round(c["gbs"], 1), frac=round(c["gbs"] * 1e9 / PEAK_BPS, 4)) for c in cases]
    for pc in per_case:
        log("  %-5s q=%d %-11s %8.4f ms  %7.1f GB/s  %.3f of peak" % (pc["layout"], pc["q"], pc["kernel"], pc["ms"],
                                                                       pc["gbs"], pc["frac"]))
    dom = max(cases, key=lambda c: c["ms"])
    achieved = dom["bytes"] / (dom["ms"] * 1e-3)
    traffic = None
    for t in (traffic_from_profiles() or {}).get("cases", []):
        if t.get("layout") == dom["layout"] and t.get("q") == dom["q"]:
            traffic = t["hbm_bytes_per_launch"]
    roofline = dict(bound="hbm", achieved=round(achieved / 1e9, 1), peak=8000.0, unit="GB/s",
                    frac=round(achieved / PEAK_BPS, 4), traffic=traffic,
                    kernel="%s (%s-order, q=%d)" % (dom["kernel"], dom["layout"], dom["q"]),
                    algorithmic_bytes_per_launch=dom["bytes"], launch_ms=round(dom["ms"], 4))

    # practical ceilings: a pure read stream over the same buffer, and the same read stream with a
    # write stream of 1/128, 1/16 and 1/4 of it (the output fractions of 128^4, 16^7 and order-8 q=4);
    # each the faster of two access patterns: one contiguous chunk per block (a persistent stream),
    # and a compact sweep of short-lived blocks (2, 4 or 8 loads per thread, the best of the three)
    partial = torch.empty(256 * 64, dtype=torch.float32, device=dev)
    fsink = torch.zeros(1, dtype=torch.float32, device=dev)
    L = ttv_amd.lib()
    wbuf = torch.empty(N // 4, dtype=torch.float32, device=dev)

    def ceiling(wbytes):
        if wbytes == 0:
            chunked = timed_cycle(torch, ttv_amd, [lambda: ttv_amd.check(L.ttvx_bench_read(
                Aflat.data_ptr(), N * 4, partial.data_ptr(), partial.numel(), stream.cuda_stream))], args.steps, sink)
        else:
            chunked = timed_cycle(torch, ttv_amd, [lambda: ttv_amd.check(L.ttvx_bench_mix(
                Aflat.data_ptr(), N * 4, wbuf.data_ptr(), wbytes, stream.cuda_stream))], args.steps, sink)
        sweeps = [timed_cycle(torch, ttv_amd, [lambda u=u: ttv_amd.check(L.ttvx_bench_sweep(
            Aflat.data_ptr(), N * 4, wbuf.data_ptr(), wbytes, u, fsink.data_ptr(), stream.cuda_stream))],
            args.steps, sink) for u in (2, 4, 8)]
        gbs = lambda ms: (N * 4 + wbytes) / (ms * 1e-3) / 1e9
        best = min([chunked] + sweeps)
        return {"gbs": round(gbs(best), 1), "frac": round(gbs(best) * 1e9 / PEAK_BPS, 4),
                "chunked_gbs": round(gbs(chunked), 1), "sweep_gbs": round(gbs(min(sweeps)), 1)}

    rc = ceiling(0)
    read_gbs = rc["gbs"]
    log("read ceiling: %.1f GB/s (%.3f of peak; chunked %.1f, sweep %.1f)" % (read_gbs, rc["frac"], rc["chunked_gbs"],
                                                                            rc["sweep_gbs"]))
    mixed = {}
    for frac in (128, 16, 4):
        mixed["write_1_%d" % frac] = m = ceiling(N * 4 // frac)
        log("read + 1/%d write ceiling: %.1f GB/s (chunked %.1f, sweep %.1f)" % (frac, m["gbs"], m["chunked_gbs"],
                                                                                m["sweep_gbs"]))
    del wbuf

    out = {
        "metric": "HBM GB/s (% peak) order-4 FP32 TTV; BF16/FP8 speedup at rel err<=1e-2",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 4),
        "ms_per_step_sum_of_cases": round(sum(c["ms"] for c in cases), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (U[-1,1], seeded)",
        "config": {"workload": "FP32 128^4 mode-q TTV, q=1..4 x {first,last}-order layouts (8 products per step)",
                   "shape": SHAPE, "elements": N, "bytes_per_product": case_bytes(128)},
        "frac_of_peak": round(value * 1e9 / PEAK_BPS, 4),
        "min_case_frac": round(min(pc["frac"] for pc in per_case), 4),
        "read_ceiling": dict(rc, value_over_ceiling=round(value / read_gbs, 4)),
        "mixed_ceilings": mixed,
        "roofline": roofline,
        "per_case": per_case,
    }

    # FP64 oracle results are needed for the precision errors; A goes to the host once
    need_host = not (args.no_precision and args.no_cpu)
    a_host = Aflat.cpu().numpy() if need_host else None


    if not args.no_precision:
        out["precision"] = precision_leg(torch, np, ttv_amd, Aflat, a_host, cases, args, log)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_leg(np, a_host, cases, args, log)
        out["host_staged"] = host_staged_leg(np, ttv_amd, a_host, cases, log)

    if not args.no_batch:
        del a_host
        out["batching"] = batching_leg(torch, np, ttv_amd, log)

    if not args.no_extra:
        out["other_configs"] = other_configs_leg(torch, np, ttv_amd, Aflat, mixed, log)

    if not args.no_order8:
        del Aflat
        for c in cases:
            c.pop("A", None)
            c.pop("C", None)
        torch.cuda.empty_cache()
        out["precision_order8"] = order8_leg(torch, np, ttv_amd, log)

    print(json.dumps(out), flush=True)


def order8_leg(torch, np, ttv_amd, log, reps=5):
    """The order-8 precision configuration: FP32 A of extents (4096,4,8,4,8,4,2,64) = 2^31 elements
    (8 GiB; the product of the extents, not the ~1.07e9 sometimes quoted for it), layout drawn from
    a seed, q in {1, 4, 8} (n_q = 4096, 4, 64); A ~ N(0,1) with 1% of entries x100, b ~ U[-1,1].
    FP32, BF16 and FP8-e4m3 storage, FP32 accumulation: mean launch time (launches back to back;
    every operand is 2-8 GiB, far above the last-level cache), speedup and its byte-bound ideal,
    relative L2 and max errors against the FP64 oracle on the original values."""
    from oracle import oracle as O
    dev = torch.device("cuda:0")
    na = [4096, 4, 8, 4, 8, 4, 2, 64]
    rng = np.random.default_rng(31)
    pia = [int(x) for x in rng.permutation(8) + 1]
    wa = O.strides(na, pia)
    n = int(np.prod(na))
    g = torch.Generator(device=dev)
    g.manual_seed(31)
    A = torch.randn(n, generator=g, device=dev)
    A.mul_(torch.where(torch.rand(n, generator=g, device=dev) < 0.01, 100.0, 1.0))
    a = A.cpu().numpy()
    Av = A.as_strided(na, wa)
    Abf = ttv_amd.quantize_bf16(A).as_strided(na, wa)
    sink = torch.zeros(1, dtype=torch.int32, device=dev)
    res = {"shape": na, "layout": pia, "elements": n, "cases": []}
    for q in (1, 4, 8):
        nq = na[q - 1]
        b = rng.uniform(-1, 1, nq).astype(np.float32)
        bd = torch.from_numpy(b).to(dev)
        ref = O.naive(q, a, na, wa, b, pia=pia)
        rn, rmax = float(np.linalg.norm(ref)), float(np.max(np.abs(ref)))
        outc = n // nq
        codes, scales, fps = ttv_amd.quantize_fp8(q, Av)
        C32 = ttv_amd.ttv(q, Av, bd)
        C16 = ttv_amd.ttv_bf16(q, Abf, bd)
        C8 = ttv_amd.ttv_fp8(q, codes, scales, fps, bd)
        runs = {"fp32": (lambda: ttv_amd.ttv(q, Av, bd, out=C32), C32, n * 4),
                "bf16": (lambda: ttv_amd.ttv_bf16(q, Abf, bd, out=C16), C16, n * 2),
                "fp8": (lambda: ttv_amd.ttv_fp8(q, codes, scales, fps, bd, out=C8), C8, n + scales.numel() * 4)}
        case = {"q": q, "n_q": nq}
        for k, (run, C, abytes) in runs.items():
            ms = timed_cycle(torch, ttv_amd, [run], reps, sink, us_per_launch=2000.0)
            nbytes = abytes + outc * 4 + nq * 4
            got = C.as_strided([C.numel()], [1]).cpu().numpy().astype(np.float64)
            d = got - ref
            case[k] = dict(ms=round(ms, 4), gbs=round(nbytes / (ms * 1e-3) / 1e9, 1), bytes=nbytes,
                           rel_l2=float(np.linalg.norm(d) / rn), max_abs=float(np.max(np.abs(d))))
        for k in ("bf16", "fp8"):
            case[k]["speedup"] = round(case["fp32"]["ms"] / case[k]["ms"], 3)
            case[k]["ideal_speedup"] = round(case["fp32"]["bytes"] / case[k]["bytes"], 3)
        # the FP32 product's ceiling: the compact-sweep stream reading the same 8 GiB and writing the
        # product's output bytes
        wbuf = torch.empty(outc, dtype=torch.float32, device=dev)
        L = ttv_amd.lib()
        fsink = torch.zeros(1, dtype=torch.float32, device=dev)
        st = torch.cuda.current_stream().cuda_stream
        cms = min(timed_cycle(torch, ttv_amd, [lambda u=u: ttv_amd.check(L.ttvx_bench_sweep(
            A.data_ptr(), n * 4, wbuf.data_ptr(), outc * 4, u, fsink.data_ptr(), st))], reps, sink,
            us_per_launch=2000.0) for u in (2, 4, 8))
        del wbuf
        case["fp32"]["ceiling_gbs"] = round((n * 4 + outc * 4) / (cms * 1e-3) / 1e9, 1)
        case["fp32"]["of_ceiling"] = round(case["fp32"]["gbs"] / case["fp32"]["ceiling_gbs"], 4)
        case["max_abs_C"] = rmax
        res["cases"].append(case)
        log("  order-8 q=%d (n_q=%d): fp32 %.3f ms %.0f GB/s (%.3f of ceiling) | bf16 x%.2f (ideal %.2f) err %.2e | "
            "fp8 x%.2f (ideal %.2f) err %.2e" % (q, nq, case["fp32"]["ms"], case["fp32"]["gbs"], case["fp32"]["of_ceiling"],
                                                 case["bf16"]["speedup"],
                          case["bf16"]["ideal_speedup"], case["bf16"]["rel_l2"], case["fp8"]["speedup"],
                          case["fp8"]["ideal_speedup"], case["fp8"]["rel_l2"]))
        del codes, scales, C32, C16, C8, ref
    return res


def host_staged_leg(np, ttv_amd, a_host, cases, log, reps=3):
    """The drop-in synchronous entry point with host buffers (what an unchanged reference caller
    does): A (1 GiB) is staged over PCIe on every call, so this rate is bounded by the link."""
    c = cases[0]
    A = np.lib.stride_tricks.as_strided(a_host, shape=SHAPE, strides=[w * 4 for w in c["wa"]])
    b = c["b"].cpu().numpy()
    ttv_amd.ttv(c["q"], A, b)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ttv_amd.ttv(c["q"], A, b)
        ts.append(time.perf_counter() - t0)
    ms = float(np.median(ts)) * 1e3
    gbs = c["bytes"] / (ms * 1e-3) / 1e9
    log("host-staged drop-in call (%s q=%d): %.2f ms, %.1f GB/s including PCIe" % (c["layout"], c["q"], ms, gbs))
    return dict(layout=c["layout"], q=c["q"], ms=round(ms, 3), gbs_including_pcie=round(gbs, 1))


def other_configs_leg(torch, np, ttv_amd, Aflat, mixed, log, reps=10):
    """The other shapes of the reference's benchmark: FP32 16^7 (2^28 elements, every q, first- and
    last-order layout; same data as 128^4; each product writes 1/16 of what it reads, so its
    ceiling is the read stream with a 1/16 write stream) and FP64 256^3 first-order with q = 2
    (128 MiB, timed over 8 copies so that every launch reads HBM)."""
    from ttv_amd.meta import generate_strides
    dev = Aflat.device
    sink = torch.zeros(1, dtype=torch.int32, device=dev)
    ceil16 = mixed["write_1_16"]["gbs"]
    res = {"fp32_16e7": [], "fp64_256e3": None, "ceiling_16e7_gbs": ceil16}
    sh = [16] * 7
    for lname, pia in (("first", list(range(1, 8))), ("last", list(range(7, 0, -1)))):
        wa = generate_strides(sh, pia)
        A = Aflat.as_strided(sh, wa)
        for q in range(1, 8):
            b = torch.rand(16, device=dev)
            C = ttv_amd.ttv(q, A, b)
            info = ttv_amd.plan_info(q, sh, wa, pia)
            nbytes = N * 4 + (N // 16) * 4 + 16 * 4
            ms = timed_cycle(torch, ttv_amd, [lambda: ttv_amd.ttv(q, A, b, out=C)], reps, sink)
            gbs = nbytes / (ms * 1e-3) / 1e9
            res["fp32_16e7"].append(dict(layout=lname, q=q, kernel=ttv_amd.variant_names()[info["variant"]],
                                         ms=round(ms, 4), gbs=round(gbs, 1), frac=round(gbs * 1e9 / PEAK_BPS, 4),
                                         of_ceiling=round(gbs / ceil16, 4)))
            log("  16^7 %-5s q=%d %-10s %.4f ms %7.1f GB/s (%.3f of the 1/16-write ceiling)" % (
                lname, q, res["fp32_16e7"][-1]["kernel"], ms, gbs, gbs / ceil16))
    n3 = 1 << 24
    A64 = Aflat[:n3].double().view(256, 256, 256).permute(2, 1, 0)     # first-order strides (1, 256, 65536)
    base = Aflat[:n3].double()
    A64s = [t.view(256, 256, 256).permute(2, 1, 0) for t in copies_for(base)]
    b64 = torch.rand(256, device=dev, dtype=torch.float64)
    C64 = ttv_amd.ttv(2, A64, b64)
    nbytes = n3 * 8 + 65536 * 8 + 256 * 8
    runs = [(lambda T=T: ttv_amd.ttv(2, T, b64, out=C64)) for T in A64s]
    ms = timed_cycle(torch, ttv_amd, runs, reps * len(runs), sink, us_per_launch=100.0)
    gbs = nbytes / (ms * 1e-3) / 1e9
    info = ttv_amd.plan_info(2, [256] * 3, list(A64.stride()), [1, 2, 3], ttv_amd.F64)
    # its ceiling: the compact-sweep read stream over 128 MiB with a 1/256 write stream (the FP64
    # product's output fraction), cycling over the same 8 copies
    L = ttv_amd.lib()
    fsink = torch.zeros(1, dtype=torch.float32, device=dev)
    wbuf = torch.empty(n3 * 8 // 256 // 4, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    cms = min(timed_cycle(torch, ttv_amd, [(lambda T=T, u=u: ttv_amd.check(L.ttvx_bench_sweep(
        T.data_ptr(), n3 * 8, wbuf.data_ptr(), wbuf.numel() * 4, u, fsink.data_ptr(), stream.cuda_stream)))
        for T in A64s], reps * len(A64s), sink, us_per_launch=100.0) for u in (2, 4, 8))
    cgbs = (n3 * 8 + wbuf.numel() * 4) / (cms * 1e-3) / 1e9
    res["fp64_256e3"] = dict(q=2, layout="first", kernel=ttv_amd.variant_names()[info["variant"]], ms=round(ms, 4),
                             gbs=round(gbs, 1), frac=round(gbs * 1e9 / PEAK_BPS, 4), copies=len(A64s),
                             ceiling_128mib_gbs=round(cgbs, 1), of_ceiling=round(gbs / cgbs, 4))
    log("  256^3 FP64 q=2 %.4f ms %7.1f GB/s (%.3f of the 128 MiB read ceiling, %.1f GB/s)" % (ms, gbs, gbs / cgbs, cgbs))
    fr = [c["frac"] for c in res["fp32_16e7"]]
    res["fp32_16e7_min_frac"] = min(fr)
    res["fp32_16e7_mean_frac"] = round(float(np.mean(fr)), 4)
    res["fp32_16e7_min_of_ceiling"] = min(c["of_ceiling"] for c in res["fp32_16e7"])
    return res


def batching_leg(torch, np, ttv_amd, log, count=512, seed=4):
    """512 independent products (order 3/4, extents U-int[16,512], <= 2^26 elements each, random q
    and layout) as one batch (one launch per tile body, replayed as a hipGraph) vs 512 sequential
    single launches with the default selection."""
    from ttv_amd.batch import Batch
    from ttv_amd.workloads import batch_problems
    specs = batch_problems(seed=seed, count=count)
    dev = torch.device("cuda:0")
    spans = [(1 + sum((n - 1) * w for n, w in zip(s["na"], s["wa"])) + 63) // 64 * 64 for s in specs]
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    pool = torch.empty(sum(spans), dtype=torch.float32, device=dev).uniform_(-1, 1, generator=g)
    probs, off = [], 0
    for s, sp in zip(specs, spans):
        A = pool[off:off + sp].as_strided(s["na"], s["wa"])
        b = torch.empty(s["na"][s["q"] - 1], dtype=torch.float32, device=dev).uniform_(-1, 1, generator=g)
        probs.append((s["q"], A, b))
        off += sp
    bt = Batch(probs)
    info = bt.info()
    seq_out = [ttv_amd.ttv(q, A, b) for q, A, b in probs]
    stream = torch.cuda.current_stream()

    sink = torch.zeros(1, dtype=torch.int32, device=dev)

    def timed(fn, reps):
        # every batch reads tens of GB: back-to-back repetitions are HBM-cold
        return timed_cycle(torch, ttv_amd, [fn], reps, sink, us_per_launch=30000.0)

    ms_graph = timed(lambda: bt.launch(graph=True), 5)
    ms_plain = timed(lambda: bt.launch(graph=False), 5)

    def sequential():
        for (q, A, b), C in zip(probs, seq_out):
            ttv_amd.ttv(q, A, b, out=C)

    ms_seq = timed(sequential, 3)
    worst = 0.0
    for Cb, Cs in zip(bt.outputs, seq_out):
        d = float(torch.linalg.vector_norm((Cb - Cs).double()) / torch.linalg.vector_norm(Cs.double()))
        worst = max(worst, d)
    gb = info["bytes"] / 1e9
    res = dict(problems=count, launches=info["launches"], blocks=info["blocks"], algorithmic_gb=round(gb, 2),
               batched_graph_ms=round(ms_graph, 3), batched_ms=round(ms_plain, 3), sequential_ms=round(ms_seq, 3),
               batched_gbs=round(gb / (ms_graph * 1e-3), 1), sequential_gbs=round(gb / (ms_seq * 1e-3), 1),
               speedup=round(ms_seq / ms_graph, 3), max_rel_diff_vs_sequential=worst)
    log("batching: %d problems %.1f GB in %d launches  graph %.3f ms (%.0f GB/s)  plain %.3f ms  sequential %.3f ms "
        "(%.0f GB/s)  speedup %.2f  max rel diff %.1e" % (count, gb, info["launches"], ms_graph, gb / (ms_graph * 1e-3),
                                                         ms_plain, ms_seq, gb / (ms_seq * 1e-3), ms_seq / ms_graph, worst))
    bt.close()
    return res


def precision_leg(torch, np, ttv_amd, Aflat, a_host, cases, args, log):
    """BF16 and FP8 storage of the 8 FP32 cases.  Each is timed like the FP32 products (launches back
    to back behind a spin, one event pair): the BF16 tensor (512 MiB) rotates over 2 copies and the
    FP8 codes (256 MiB) over 4, so every launch reads A from HBM like the 1 GiB FP32 tensor does;
    the speedup divides the FP32 case's mean launch time by this one."""
    from oracle import oracle as O
    res = {"bf16": [], "fp8": []}
    sink = torch.zeros(1, dtype=torch.int32, device=Aflat.device)
    Abf = ttv_amd.quantize_bf16(Aflat)            # same strides (flat)
    Abfs = copies_for(Abf)
    abf_host = None
    reps = max(12, args.steps)
    for c in cases:
        q, wa, pia = c["q"], c["wa"], c["pia"]
        b = c["b"]
        bh = b.cpu().numpy()
        ref = O.naive(q, a_host, SHAPE, wa, bh, pia=pia)
        rn = np.linalg.norm(ref)
        c32 = c["C"].as_strided([c["C"].numel()], [1]).cpu().numpy().astype(np.float64)
        err32 = float(np.linalg.norm(c32 - ref) / rn)
        for fmt in ("bf16", "fp8"):
            if fmt == "bf16":
                Cq = ttv_amd.ttv_bf16(q, Abf.as_strided(SHAPE, wa), b)
                runs = [(lambda T=T.as_strided(SHAPE, wa): ttv_amd.ttv_bf16(q, T, b, out=Cq)) for T in Abfs]
                extra = 0
                elem = 2
            else:
                codes, scales, fps = ttv_amd.quantize_fp8(q, Aflat.as_strided(SHAPE, wa))
                Cq = ttv_amd.ttv_fp8(q, codes, scales, fps, b)
                flat = codes.as_strided([N], [1])
                runs = [(lambda T=T.as_strided(SHAPE, wa): ttv_amd.ttv_fp8(q, T, scales, fps, b, out=Cq))
                        for T in copies_for(flat)]
                extra = scales.numel() * 4
                elem = 1
            # best of three regions, like the FP32 per-case times (a region can catch a stall of the box)
            ms = min(timed_cycle(torch, ttv_amd, runs, reps, sink, us_per_launch=150.0) for _ in range(3))
            got = Cq.as_strided([Cq.numel()], [1]).cpu().numpy().astype(np.float64)
            err = float(np.linalg.norm(got - ref) / rn)
            # error of the accumulation alone: FP64 oracle on the dequantised values
            if fmt == "bf16":
                if abf_host is None:
                    abf_host = O.bf16_to_f32(Abf.cpu().numpy().view(np.uint16))
                ref_q = O.naive(q, abf_host, SHAPE, wa, bh, pia=pia)
            else:
                # scales are constant along each fiber: C = scale[j // fps] * sum_k decode(code) b_k
                table = O.e4m3_decode(np.arange(256, dtype=np.uint8))
                deq = table[flat.cpu().numpy()]
                sh = scales.cpu().numpy().astype(np.float64)
                ref_q = O.naive(q, deq, SHAPE, wa, bh, pia=pia)
                ref_q *= sh[np.arange(ref_q.size) // fps]
                del deq
            err_acc = float(np.linalg.norm(got - ref_q) / np.linalg.norm(ref_q))
            nbytes = case_bytes(128, elem=elem, extra=extra)
            res[fmt].append(dict(layout=c["layout"], q=q, ms=round(ms, 4), speedup=round(c["ms"] / ms, 3),
                                 gbs=round(nbytes / (ms * 1e-3) / 1e9, 1), frac=round(nbytes / (ms * 1e-3) / PEAK_BPS, 4),
                                 ideal_speedup=round(c["bytes"] / nbytes, 3),
                                 rel_err_vs_fp64=err, rel_err_vs_fp64_dequantised=err_acc, fp32_rel_err=err32))
            log("  %s %-5s q=%d %.4f ms  speedup %.2f (ideal %.2f)  rel err %.2e (accumulation %.2e)" % (
                fmt, c["layout"], q, ms, c["ms"] / ms, c["bytes"] / nbytes, err, err_acc))
            del runs, Cq
            if fmt == "fp8":
                del codes, scales, flat
    summ = {}
    for fmt in ("bf16", "fp8"):
        sp = [r["speedup"] for r in res[fmt]]
        summ[fmt] = dict(speedup_geomean=round(float(np.exp(np.mean(np.log(sp)))), 3),
                         speedup_min=round(min(sp), 3),
                         rel_err_max=max(r["rel_err_vs_fp64"] for r in res[fmt]),
                         rel_err_dequantised_max=max(r["rel_err_vs_fp64_dequantised"] for r in res[fmt]),
                         per_case=res[fmt])
    return summ


def cpu_leg(np, a_host, cases, args, log):
    from oracle import oracle as O
    blas = O.load_blas()
    cores = O.number_cores()
    t_total, nbytes, reps = 0.0, 0, 0
    bh = [c["b"].cpu().numpy() for c in cases]
    C = [np.zeros(N // 128, np.float32) for _ in cases]
    # warm-up of every case, then whole steps until the time budget is used (at least one step)
    for i, c in enumerate(cases):
        O.log_cpu(c["q"], a_host, SHAPE, c["wa"], c["pia"], bh[i], c_out=C[i], threads=cores)
    while reps < 1 or t_total < args.cpu_seconds:
        t0 = time.perf_counter()
        for i, c in enumerate(cases):
            O.log_cpu(c["q"], a_host, SHAPE, c["wa"], c["pia"], bh[i], c_out=C[i], threads=cores)
        t_total += time.perf_counter() - t0
        nbytes += sum(c["bytes"] for c in cases)
        reps += 1
    gbs = nbytes / t_total / 1e9
    log("cpu baseline: %.2f GB/s on %d cores (%s), %d steps in %.1f s" % (gbs, cores, blas or "no BLAS", reps, t_total))
    return {"value": round(gbs, 2), "unit": "GB/s", "cores": cores, "kind": "port",
            "blas": os.path.basename(blas) if blas else "none (reference's own OpenMP loops)",
            "sample": "%d full steps (8 products of FP32 128^4, same inputs) in %.1f s" % (reps, t_total)}


if __name__ == "__main__":
    main()
