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
the stream the kernels run on; `cold_launch` is the forward's first pass (input
read from HBM, rotated buffers) and `warm_launch` the dominant kernel's inverse
launches (input just written, largely cache-resident).  `cpu_baseline` times the
FP64 CPU oracle on the host cores (all OpenMP threads) on the same workload
shape.

The other measured configurations are reported under their own keys (forward
transforms, default plans, transform-level %HBM = algorithmic bytes of the
whole transform / its time, forward error vs the FP64 oracle on a sample and
forward+inverse error in FP64 on the GPU over the whole batch):
  c1  1D N=256 x 4096, re/im U[-1,1)
  c3  2D 4096 x 4096 x 4, re/im U[-1,1) + 100 exp(2 pi i (kx m + ky n)/4096) at a
      random (kx, ky) per transform; FP16, BF16 and split
  c4  3D 512^3, re/im N(0,1)
  c6  1D N=65536 x 4096, re/im U[-1,1)
  c5  the 20064-item ragged batch (20000 1D N = 2^8..2^18 log-uniform + 64 2D
      with sides in {256..2048}) in its drawn order: one ragged exec against
      per-shape batched launches (items grouped by shape) and one launch per
      item (drawn order); plain and split
`--no-configs` skips them.
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
    ap.add_argument("--no-configs", action="store_true", help="skip the c1/c3/c4/c5/c6 legs")
    ap.add_argument("--configs", default="c1,c3,c4,c6,c5", help="which of c1,c3,c4,c5,c6 to run")
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


ELEM_BYTES = {"fp16": 4, "bf16": 4, "split": 8}


def timed_ms(torch, stream, fn, steps, min_warm_ms=100.0, warmup=2):
    """Mean ms per call over `steps` calls after at least `warmup` calls and
    `min_warm_ms` of GPU time untimed (events on `stream`)."""
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    spent, i, n = 0.0, 0, warmup
    while True:
        e0.record(stream)
        for _ in range(n):
            fn(i)
            i += 1
        e1.record(stream)
        torch.cuda.synchronize()
        spent += e0.elapsed_time(e1)
        if spent >= min_warm_ms:
            break
        n = min(256, 2 * n)
    e0.record(stream)
    for k in range(steps):
        fn(k)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def roundtrip_rel_l2(mc, y, fmt, exps, n, ref):
    """Worst rel-L2 of inv(fwd(x)) / n against x over runs of transforms that
    share an exponent (FP64 reduction on the GPU); exps: per-transform output
    exponents of the inverse, y its output (n elements per transform)."""
    worst, t = 0.0, 0
    while t < len(exps):
        u = t
        while u < len(exps) and exps[u] == exps[t]:
            u += 1
        rel, _ = mc.error_norms(y[t * n:u * n], fmt, int(exps[t]) - int(math.log2(n)), ref[2 * t * n:2 * u * n])
        worst = max(worst, rel)
        t = u
    return worst


def c3_field(torch, batch, seed=3):
    """4096 x 4096 x batch: re/im U[-1,1) + 100 exp(2 pi i (kx m + ky n)/4096) at
    a random (kx, ky) per transform, as (elements, 2) float32 on the device."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.rand((batch, 4096, 4096, 2), device="cuda", generator=g) * 2 - 1
    rng = np.random.default_rng(seed)
    m = torch.arange(4096, device="cuda", dtype=torch.int64)
    for b in range(batch):
        kx, ky = (int(v) for v in rng.integers(0, 4096, 2))
        ph = ((kx * m[:, None] + ky * m[None, :]) % 4096).to(torch.float64) * (2 * math.pi / 4096)
        x[b, :, :, 0] += (100.0 * torch.cos(ph)).float()
        x[b, :, :, 1] += (100.0 * torch.sin(ph)).float()
    return x.reshape(-1, 2)


def forward_leg(torch, mc, oracle, stream, dims, batch, x32, fi, fo, steps, check):
    """Forward transforms with the default plan: time, transform-level %HBM,
    forward error of the first `check` transforms vs the FP64 oracle and the
    forward+inverse error over the whole batch (FP64 on the GPU)."""
    n = int(np.prod(dims))
    elems = n * batch
    d_in = x32.bfloat16() if fi == "bf16" else x32.half()
    rot = max(1, min(3, (2 * MALL_BYTES) // (elems * ELEM_BYTES[fi]) + 1))
    ins = [d_in] + [d_in.clone() for _ in range(rot - 1)]
    d_out = mc.empty(elems, fo)
    p = mc.Plan(dims, batch, fi, fo)
    st = stream.cuda_stream
    in_ptrs, out_ptr = [t.data_ptr() for t in ins], d_out.data_ptr()
    ms = timed_ms(torch, stream, lambda i: p.exec_ptr(in_ptrs[i % rot], out_ptr, mc.FORWARD, st), steps)
    flops = 5.0 * n * math.log2(n) * batch
    alg = elems * (ELEM_BYTES[fi] + ELEM_BYTES[fo])
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    flags = p.status_flags()
    e_fwd = p.exponents()
    res = {"in": fi, "out": fo, "ms": round(ms, 4), "gflops": round(flops / (ms * 1e-3) / 1e9, 1),
           "hbm_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "alg_bytes": alg,
           "exponents": [int(e_fwd.min()), int(e_fwd.max())]}
    k = min(batch, check)
    if k > 0:
        y = p.result(d_out, k)
        ref = oracle.fft(mc.from_device(d_in[: k * n], fi), tuple(dims))
        res.update({"rel_l2_fwd": oracle.rel_l2(y, ref), "max_abs_fwd": oracle.max_abs(y, ref), "checked_fwd": k})
        del ref, y
    del ins
    # forward + inverse over the whole batch, checked on the GPU in FP64
    mid_fmt = fo
    pi = mc.Plan(dims, batch, mid_fmt, mid_fmt)
    d_rt = mc.empty(elems, mid_fmt)
    pi.exec(d_out, d_rt, mc.INVERSE, in_exp=p)
    ref64 = (d_in.float() if fi == "bf16" else d_in).double().reshape(-1)
    res["rel_l2_roundtrip"] = roundtrip_rel_l2(mc, d_rt, mid_fmt, pi.exponents(), n, ref64)
    res["status_flags"] = flags | pi.status_flags()
    del ref64, d_rt, d_out, d_in
    p.destroy()
    pi.destroy()
    return res


def c5_leg(torch, mc, oracle, stream, fi, fo, steps):
    """The 20064-item ragged batch in drawn order (items back to back as drawn):
    one ragged exec against per-shape batched launches over the same items
    grouped by shape and against one launch per item."""
    rng = np.random.default_rng(5)
    shapes = [[int(2 ** k)] for k in rng.integers(8, 19, 20000)]
    sides = [256, 512, 1024, 2048]
    shapes += [[int(rng.choice(sides)), int(rng.choice(sides))] for _ in range(64)]
    sizes = np.array([int(np.prod(s)) for s in shapes], dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sizes.sum())
    flops = float(sum(5.0 * s * math.log2(s) for s in sizes))
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    d_in = torch.randn((total, 2), dtype=torch.float16, device="cuda", generator=g)
    d_out = mc.empty(total, fo)
    st = stream.cuda_stream
    eb_in, eb_out = ELEM_BYTES[fi], ELEM_BYTES[fo]
    rp = mc.RaggedPlan([(s, int(o), int(o)) for s, o in zip(shapes, offs)], fi, fo)
    ms_r = timed_ms(torch, stream, lambda i: rp.exec(d_in, d_out, mc.FORWARD), steps)
    flags = rp.status_flags()
    # per-shape batched launches: the same items grouped by shape (a sorted copy)
    order = sorted(range(len(shapes)), key=lambda i: (len(shapes[i]), shapes[i]))
    s_in = torch.empty_like(d_in)
    s_out = mc.empty(total, fo)
    o = 0
    groups = []
    for i in order:
        n = int(sizes[i])
        s_in[o:o + n] = d_in[int(offs[i]):int(offs[i]) + n]
        if groups and groups[-1][0] == tuple(shapes[i]):
            groups[-1][2] += 1
        else:
            groups.append([tuple(shapes[i]), o, 1])
        o += n
    plans = [(mc.Plan(list(s), cnt, fi, fo, "balanced"), o0) for s, o0, cnt in groups]
    ptrs = [(p, s_in.data_ptr() + o0 * eb_in, s_out.data_ptr() + o0 * eb_out) for p, o0 in plans]

    def per_shape(i):
        for p, a, b in ptrs:
            p.exec_ptr(a, b, mc.FORWARD, st)

    ms_p = timed_ms(torch, stream, per_shape, steps)
    # one launch per item in drawn order (a batch-1 plan per shape)
    single = {tuple(s): mc.Plan(list(s), 1, fi, fo, "balanced") for s, _, _ in groups}
    items = [(single[tuple(s)], d_in.data_ptr() + int(o0) * eb_in, d_out.data_ptr() + int(o0) * eb_out)
             for s, o0 in zip(shapes, offs)]

    def per_item(i):
        for p, a, b in items:
            p.exec_ptr(a, b, mc.FORWARD, st)

    ms_i = timed_ms(torch, stream, per_item, 2, min_warm_ms=0.0, warmup=1)
    # accuracy of the ragged exec on a sample of items
    rp.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    flags |= rp.status_flags()
    worst = 0.0
    for i in rng.choice(len(shapes), 24, replace=False):
        o0, n = int(offs[i]), int(sizes[i])
        y = mc.from_device(d_out[o0:o0 + n], fo, rp.exponents[i])
        worst = max(worst, oracle.rel_l2(y, oracle.fft(mc.from_device(d_in[o0:o0 + n], fi), tuple(shapes[i]))))
    alg = total * (eb_in + eb_out)
    res = {"in": fi, "out": fo, "items": len(shapes), "elements": total, "order": "drawn",
           "ragged_ms": round(ms_r, 4), "ragged_gflops": round(flops / (ms_r * 1e-3) / 1e9, 1),
           "ragged_hbm_frac": round(alg / (ms_r * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "per_shape_launches": len(plans), "per_shape_ms": round(ms_p, 4),
           "speedup_vs_per_shape": round(ms_p / ms_r, 3),
           "per_item_launches": len(items), "per_item_ms": round(ms_i, 4),
           "speedup_vs_per_item": round(ms_i / ms_r, 3),
           "worst_rel_l2_sample": worst, "status_flags": flags}
    for p, _ in plans:
        p.destroy()
    for p in single.values():
        p.destroy()
    rp.destroy()
    return res


def run_configs(torch, mc, oracle, stream, args):
    steps = max(2, min(args.steps, 10))
    want = set(args.configs.split(","))
    out = {}

    def note(msg):
        print(f"bench: {msg}", file=sys.stderr, flush=True)

    if "c1" in want:
        g = torch.Generator(device="cuda")
        g.manual_seed(1)
        x = torch.rand((256 * 4096, 2), device="cuda", generator=g) * 2 - 1
        out["c1"] = {"workload": "1D N=256 x 4096 forward, re/im U[-1,1)",
                     "legs": [forward_leg(torch, mc, oracle, stream, [256], 4096, x, fi, fo, steps, 64)
                              for fi, fo in (("fp16", "fp16"), ("fp16", "split"))]}
        note("c1 done")
    if "c3" in want:
        x = c3_field(torch, 4)
        out["c3"] = {"workload": "2D 4096x4096 x 4 forward, re/im U[-1,1) + amplitude-100 2D exponential "
                                 "at a random (kx, ky) per transform (seed 3)",
                     "legs": [forward_leg(torch, mc, oracle, stream, [4096, 4096], 4, x, fi, fo, steps, 1)
                              for fi, fo in (("fp16", "fp16"), ("bf16", "bf16"), ("fp16", "split"))]}
        del x
        note("c3 done")
    if "c4" in want:
        g = torch.Generator(device="cuda")
        g.manual_seed(4)
        x = torch.randn((512 ** 3, 2), device="cuda", generator=g)
        out["c4"] = {"workload": "3D 512^3 x 1 forward, re/im N(0,1)",
                     "legs": [forward_leg(torch, mc, oracle, stream, [512, 512, 512], 1, x, fi, fo, steps, 0)
                              for fi, fo in (("fp16", "fp16"), ("fp16", "split"))]}
        del x
        note("c4 done")
    if "c6" in want:
        g = torch.Generator(device="cuda")
        g.manual_seed(6)
        x = torch.rand((65536 * 4096, 2), device="cuda", generator=g) * 2 - 1
        out["c6"] = {"workload": "1D N=65536 x 4096 forward, re/im U[-1,1)",
                     "legs": [forward_leg(torch, mc, oracle, stream, [65536], 4096, x, fi, fo, steps, 8)
                              for fi, fo in (("fp16", "fp16"), ("fp16", "split"))]}
        del x
        note("c6 done")
    if "c5" in want:
        out["c5"] = {"workload": "ragged: 20000 1D N=2^8..2^18 log-uniform + 64 2D sides {256..2048}, "
                                 "re/im N(0,1), drawn order",
                     "legs": [c5_leg(torch, mc, oracle, stream, fi, fo, steps)
                              for fi, fo in (("fp16", "fp16"), ("fp16", "split"))]}
        note("c5 done")
    torch.cuda.empty_cache()
    return out


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
    launches = []   # (kernel name, pass, direction, start event, stop event)
    for i in range(args.warmup + args.steps):
        if i == args.warmup:
            s0.record(stream)
        for j, (src, dst, d) in enumerate(((ins[i % n_rot], mid, mc.FORWARD), (mid, out, mc.INVERSE))):
            starts, stops = ev[2 * i + j]
            pf.exec_timed(src, dst, d, starts, stops)
            if i >= args.warmup:
                launches += [(names[p_i], p_i, j, starts[p_i], stops[p_i]) for p_i in range(n_pass)]
    s1.record(stream)
    torch.cuda.synchronize()
    ms_timed_step = s0.elapsed_time(s1) / args.steps
    per, per_dir = {}, {}
    for name, p_i, j, e0, e1 in launches:
        t = e0.elapsed_time(e1)
        per.setdefault(name, []).append(t)
        per_dir.setdefault((p_i, j), []).append(t)
    kernels = []
    for p_i, name in enumerate(names):
        ms_k = sum(per[name]) / len(per[name])
        fwd, inv = per_dir[(p_i, 0)], per_dir[(p_i, 1)]
        kernels.append({"pass": p_i, "kernel": name, "ms": ms_k, "launches": len(per[name]),
                        "ms_forward": sum(fwd) / len(fwd), "ms_inverse": sum(inv) / len(inv),
                        "alg_gbs": 8.0 * elems / (ms_k * 1e-3) / 1e9})
    dom = max(kernels, key=lambda r: r["ms"])
    traffic = traffic_from_profiles(dom["kernel"])

    def launch_rate(ms_k):
        gbs = 8.0 * elems / (ms_k * 1e-3) / 1e9
        return {"ms": round(ms_k, 4), "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}

    roofline = {"bound": "hbm", "achieved": round(dom["alg_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dom["alg_gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": dom["kernel"], "kernel_ms": round(dom["ms"], 4),
                "alg_bytes_per_launch": 8 * elems,
                "cold_launch": dict(kernel=kernels[0]["kernel"], direction="forward",
                                    **launch_rate(kernels[0]["ms_forward"])),
                "warm_launch": dict(kernel=dom["kernel"], direction="inverse", **launch_rate(dom["ms_inverse"]))}
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

    configs = None if args.no_configs else run_configs(torch, mc, oracle, stream, args)

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
    if configs:
        line.update(configs)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
