#!/usr/bin/env python3
"""Benchmark of the batched C2C FFT hot path on one MI355X.

Workload (one "step"): 1D complex-to-complex FFT, N = 131072 (16^4 x 2, the
radix-2 tail case), batch 256, forward then inverse, interleaved complex FP16,
re/im iid N(0,1) from a seed -- plain FP16 (the `value` line) and split FP16
hi+lo (reported beside it), each with its rel-L2 / max-abs error against an
FP64 FFT of the same quantised input.

`value` is GFLOP/s with the 5 N log2 N convention over the forward + inverse
transforms of the whole batch, inputs resident in HBM.  Inputs rotate over
several buffers (> 2x the 256 MB Infinity Cache) so the timed region reads HBM.
Warm-up is at least `--warmup` steps and at least `--warmup-ms` of GPU time
(default 100 ms): the GPU's clocks take about 40 ms of sustained load to settle
(a step from idle takes ~12 % longer), and every timed leg, rocFFT's included,
is measured after the same warm-up.

`roofline` is for the dominant kernel of the plain path: algorithmic bytes per
launch (8 B per complex element: one FP16 read + one FP16 write, for every
element the launch covers) / its average duration, measured with HIP events on
the stream the kernels run on.  `cpu_baseline` times the FP64 CPU oracle on the
host cores (all OpenMP threads) on the same workload shape.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, BATCH = 131072, 256
HBM_PEAK_GBS = 8000.0
MALL_BYTES = 256 << 20


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--warmup-ms", type=float, default=100.0,
                    help="minimum GPU time of the untimed warm-up before each timed leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU baseline sample time")
    ap.add_argument("--no-split", action="store_true", help="skip the split-precision leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy 5 times (scales rocprofv3 byte counters)")
    return ap.parse_args()


def traffic_from_profiles(kernel_name: str):
    """HBM bytes per launch of `kernel_name` from the committed rocprofv3 PMC
    summary (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        e = d.get("kernels", {}).get(kernel_name)
        return None if e is None else float(e["hbm_bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); multi-GPU scaling is not part "
              "of this build.", file=sys.stderr)
        sys.exit(2)
    import torch

    import mcfft_amd as mc

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        a = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": a.numel() * 4, "launches": 5}))
        return
    flops_step = 2 * 5.0 * N * math.log2(N) * BATCH   # forward + inverse
    elems = N * BATCH

    rng = np.random.default_rng(2)
    x = rng.standard_normal(elems) + 1j * rng.standard_normal(elems)
    xq = mc.decode(mc.encode(x, "fp16"), "fp16")
    # rotate inputs: each copy 128 MiB; together > 2x the Infinity Cache
    n_rot = max(2, (2 * MALL_BYTES) // (elems * 4) + 1)
    base = mc.to_device(x, "fp16")
    ins = [base] + [base.clone() for _ in range(n_rot - 1)]
    mid = mc.empty(elems, "fp16")
    out = mc.empty(elems, "fp16")

    # the default plan: block floating point (per-transform exponents found on
    # the GPU); the inverse takes the forward's exponents as its input exponents
    pf = mc.Plan([N], BATCH, "fp16", "fp16")
    stream = torch.cuda.current_stream()

    def step(i):
        pf.exec(ins[i % n_rot], mid, mc.FORWARD)
        pf.exec(mid, out, mc.INVERSE, in_exp=pf)

    def warm(fn, warmup, min_ms):
        """Untimed: `warmup` calls, then more (in groups of 10) until at least
        `min_ms` of GPU time has passed."""
        w0 = torch.cuda.Event(enable_timing=True)
        w1 = torch.cuda.Event(enable_timing=True)
        w0.record(stream)
        for i in range(warmup):
            fn(i)
        w1.record(stream)
        torch.cuda.synchronize()
        spent, i = w0.elapsed_time(w1), warmup
        while spent < min_ms:
            w0.record(stream)
            for _ in range(10):
                fn(i)
                i += 1
            w1.record(stream)
            torch.cuda.synchronize()
            spent += w0.elapsed_time(w1)

    def timed(fn, steps, warmup, per_step=None, min_warm_ms=None):
        """Mean ms per step over `steps` steps between two events, after the
        warm-up above; with `per_step` (a list), also the duration of every step
        (an event at each step boundary)."""
        warm(fn, warmup, args.warmup_ms if min_warm_ms is None else min_warm_ms)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        ev[0].record(stream)
        for i in range(steps):
            fn(i)
            if per_step is not None or i == steps - 1:
                ev[i + 1].record(stream)
        torch.cuda.synchronize()
        if per_step is not None:
            per_step.extend(ev[i].elapsed_time(ev[i + 1]) for i in range(steps))
        return ev[0].elapsed_time(ev[steps]) / steps

    step_ms = []
    ms_plain = timed(step, args.steps, args.warmup, step_ms)

    # errors (plain): forward output of the first 8 transforms vs FP64, and the round trip
    import oracle
    k = 8 * N
    pf.exec(ins[0], mid, mc.FORWARD)
    y_f = pf.result(mid, 8)
    pf.exec(mid, out, mc.INVERSE, in_exp=pf)
    y_rt = pf.result(out, 8) / N
    ref_f = oracle.fft(xq[:k], (N,))
    plain_err = {"rel_l2_fwd": oracle.rel_l2(y_f, ref_f), "max_abs_fwd": oracle.max_abs(y_f, ref_f),
                 "rel_l2_roundtrip": oracle.rel_l2(y_rt, xq[:k]), "checked_transforms": 8}

    # per-kernel durations of the plain path: the same forward+inverse steps, each
    # pass launched with a start/stop HIP event pair that the runtime attaches to
    # the kernel dispatch itself (mcfft_exec_timed), on the launch stream;
    # averaged per kernel over its forward and inverse launches (as rocprofv3
    # --stats does).  The events are created before the timed region.
    n_pass = pf.pass_count()
    desc = pf.describe().strip().splitlines()
    names = [desc[1 + p_i].split()[1] for p_i in range(n_pass)]
    ev = [[[torch.cuda.Event(enable_timing=True) for _ in range(n_pass)] for _ in range(2)]
          for _ in range(2 * (args.warmup + args.steps))]
    for group in ev:
        for row in group:
            for e in row:
                e.record(stream)    # creates the event
    torch.cuda.synchronize()
    warm(step, 0, args.warmup_ms)   # the error check above left the GPU idle
    s0 = torch.cuda.Event(enable_timing=True)
    s1 = torch.cuda.Event(enable_timing=True)
    launches = []   # (kernel name, start event, stop event)
    for i in range(args.warmup + args.steps):
        if i == args.warmup:
            s0.record(stream)
        for j, (src, dst, d) in enumerate(((ins[i % n_rot], mid, mc.FORWARD), (mid, out, mc.INVERSE))):
            starts, stops = ev[2 * i + j]
            pf.exec_timed(src, dst, d, starts, stops)
            if i >= args.warmup:
                launches += [(names[p_i], starts[p_i], stops[p_i]) for p_i in range(n_pass)]
    s1.record(stream)
    torch.cuda.synchronize()
    ms_timed_step = s0.elapsed_time(s1) / args.steps
    per = {}
    for name, e0, e1 in launches:
        per.setdefault(name, []).append(e0.elapsed_time(e1))
    kernels = []
    for p_i, name in enumerate(names):
        ms_k = sum(per[name]) / len(per[name])
        kernels.append({"pass": p_i, "kernel": name, "ms": ms_k, "launches": len(per[name]),
                        "alg_gbs": 8.0 * elems / (ms_k * 1e-3) / 1e9})
    dom = max(kernels, key=lambda r: r["ms"])
    traffic = traffic_from_profiles(dom["kernel"])
    roofline = {"bound": "hbm", "achieved": round(dom["alg_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dom["alg_gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": dom["kernel"], "kernel_ms": round(dom["ms"], 4),
                "alg_bytes_per_launch": 8 * elems}
    # MFMA work of the same kernel: 128 flop per complex element per radix-16
    # stage (two 32x32x16 MFMAs per 16-point DFT of 32 columns); the C2 passes
    # have two radix-16 stages each (256 = 16^2, 512 = 16^2 x 2 with a VALU
    # radix-2 tail).  rocprofv3's SQ_INSTS_VALU_MFMA_MOPS_F16 x 512 gives the
    # same count per launch.
    mfma_flops = 128.0 * 2 * elems
    mfma = {"flops_per_launch": mfma_flops, "achieved_tflops": round(mfma_flops / (dom["ms"] * 1e-3) / 1e12, 1),
            "peak_tflops": 2500.0, "frac": round(mfma_flops / (dom["ms"] * 1e-3) / 1e12 / 2500.0, 4),
            "kernel": dom["kernel"]}

    # the same step with static BALANCED scaling (one exponent for every
    # transform): the cost of the data-following scaling is the difference
    pb = mc.Plan([N], BATCH, "fp16", "fp16", "balanced")

    def bstep(i):
        pb.exec(ins[i % n_rot], mid, mc.FORWARD)
        pb.exec(mid, out, mc.INVERSE)

    ms_balanced = timed(bstep, args.steps, args.warmup)
    pb.destroy()

    # split precision leg: forward FP16 -> split, inverse split -> split
    split = None
    if not args.no_split:
        ps_f = mc.Plan([N], BATCH, "fp16", "split")
        ps_i = mc.Plan([N], BATCH, "split", "split")
        smid = mc.empty(elems, "split")
        sout = mc.empty(elems, "split")

        def sstep(i):
            ps_f.exec(ins[i % n_rot], smid, mc.FORWARD)
            ps_i.exec(smid, sout, mc.INVERSE, in_exp=ps_f)

        ms_split = timed(sstep, args.steps, args.warmup)
        ps_f.exec(ins[0], smid, mc.FORWARD)
        ps_i.exec(smid, sout, mc.INVERSE, in_exp=ps_f)
        ys = ps_f.result(smid, 8)
        ys_rt = ps_i.result(sout, 8) / N
        split_bytes = elems * (12 + 16)   # fwd FP16 in + hi/lo out, inv hi/lo in + out
        split = {"value": round(flops_step / (ms_split * 1e-3) / 1e9, 1), "ms_per_step": round(ms_split, 4),
                 "hbm_frac_job": round(split_bytes / (ms_split * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                 "rel_l2_fwd": oracle.rel_l2(ys, ref_f), "max_abs_fwd": oracle.max_abs(ys, ref_f),
                 "rel_l2_roundtrip": oracle.rel_l2(ys_rt, xq[:k])}
        del smid, sout
        ps_f.destroy()
        ps_i.destroy()

    # CPU baseline: the FP64 oracle on the host cores, same shape, bounded sample
    cpu = None
    if not args.no_cpu:
        n_tr = 16
        sample = x[: n_tr * N]
        t0 = time.perf_counter()
        yy = oracle.fft(sample, (N,))
        oracle.fft(yy, (N,), sign=+1)
        t1 = time.perf_counter() - t0
        reps = max(1, int(args.cpu_seconds / max(t1, 1e-3)))
        reps = min(reps, 64)
        t0 = time.perf_counter()
        for _ in range(reps):
            yy = oracle.fft(sample, (N,))
            oracle.fft(yy, (N,), sign=+1)
        tc = (time.perf_counter() - t0) / reps
        cpu_gflops = 2 * 5.0 * N * math.log2(N) * n_tr / tc / 1e9
        cpu = {"value": round(cpu_gflops, 2), "unit": "GFLOP/s", "cores": oracle.num_threads(), "kind": "port",
               "sample": f"FP64 forward+inverse of {n_tr} of the {BATCH} transforms (N={N}), {reps} repetitions, "
                         f"{reps * tc:.1f} s"}

    # same-run device copy bandwidth (the practical HBM ceiling) and the vendor
    # library (rocFFT, half precision, same data and step) for comparison
    ca = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
    cb = torch.empty_like(ca)
    ms_copy = timed(lambda i: cb.copy_(ca), 10, 2, min_warm_ms=0.0)
    copy_gbs = 2.0 * ca.numel() / (ms_copy * 1e-3) / 1e9
    del ca, cb
    vendor = None
    try:
        from mcfft_amd import vendor_baseline as vb
        rf = vb.RocfftPlan([N], BATCH, "half")
        ri = vb.RocfftPlan([N], BATCH, "half", inverse=True)
        st = stream.cuda_stream

        def vstep(i):
            rf.exec(ins[i % n_rot].data_ptr(), mid.data_ptr(), st)
            ri.exec(mid.data_ptr(), out.data_ptr(), st)

        ms_v = timed(vstep, args.steps, args.warmup)
        vendor = {"library": "rocFFT", "precision": "half", "ms_per_step": round(ms_v, 4),
                  "gflops": round(flops_step / (ms_v * 1e-3) / 1e9, 1), "speedup": round(ms_v / ms_plain, 3)}
        rf.destroy()
        ri.destroy()
    except Exception as e:  # noqa: BLE001 -- a missing vendor library is reported, not fatal
        vendor = {"library": "rocFFT", "error": str(e)}

    value = flops_step / (ms_plain * 1e-3) / 1e9
    job_bytes = 2 * 8.0 * elems
    line = {
        "metric": "C2C GFLOP/s (5NlogN) & %HBM roofline, split vs plain FP16, rel-L2 vs FP64",
        "value": round(value, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_ms_min": args.warmup_ms,
        "ms_per_step": round(ms_plain, 4),
        "ms_per_step_median": round(float(np.median(step_ms)), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16",
        "data": "synthetic",
        "config": {"workload": "1D C2C forward+inverse, N=131072 (16^4 x 2), batch 256, interleaved complex FP16, "
                               "re/im iid N(0,1) seed 2",
                   "n": N, "batch": BATCH, "precision": "plain fp16 (value); split fp16 hi+lo in 'split'",
                   "scaling": "auto (block floating point, per-transform exponents; the default plan)"},
        "static_balanced": {"ms_per_step": round(ms_balanced, 4),
                            "value": round(flops_step / (ms_balanced * 1e-3) / 1e9, 1)},
        "hbm_frac_job": round(job_bytes / (ms_plain * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "gflops_12nlogn": round(value * 12.0 / 5.0, 1),   # the reference's FLOP convention (test_1d.cpp)
        "copy_gbs": round(copy_gbs, 1),
        "vendor_baseline": vendor,
        "plain": plain_err,
        "split": split,
        "kernels": kernels,
        "ms_per_step_kernel_timed": round(ms_timed_step, 4),   # the steps of the per-kernel timing loop
        "roofline": roofline,
        "mfma": mfma,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
