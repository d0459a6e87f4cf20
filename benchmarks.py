"""Throughput and accuracy of every workload the library is measured on.

    python benchmarks.py [--config c1,c2,c3,c4,c5,c6,ragged_small,real,layout] [--reps 10]

One JSON line per (config, precision): GFLOP/s (5 N log2 N, N = elements per
transform), fraction of the 8 TB/s HBM roofline computed from algorithmic bytes
(one read of the input + one write of the output in their storage formats),
and rel-L2 / max-abs error against an FP64 FFT of the same quantised input
(computed on a sample of transforms where the full batch would take long on the
host).  `c5` is the 20064-item ragged batch (about 1e9 elements) as phased
ragged launches, as one launch with in-kernel dependencies and as per-shape
batched launches; `ragged_small` compares the persistent ragged launches with
per-shape batched launches (eagerly and captured in a hipGraph) and with one
launch per transform, over the same data; `real` is the 1D R2C path and `layout` the
planar-layout exec.

The FP64 reference used for the error columns is the repository's CPU oracle
(oracle/), imported only here, as the checker.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import mcfft_amd as mc  # noqa: E402

HBM = 8000.0  # GB/s
MALL = 256 << 20


def _timer(torch, fn, reps, warmup=2, min_ms=2.0, warm_ms=60.0):
    """Mean time per call of `reps` back-to-back calls, with the rep count raised
    until the timed region lasts at least `min_ms` (a few-microsecond exec is
    otherwise dominated by the jitter of its first launches).  Before timing,
    at least `warmup` calls and `warm_ms` of GPU time run untimed: the GPU's
    clocks settle after about 40 ms of sustained load."""
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    spent, n, i = 0.0, warmup, 0
    while True:
        e0.record()
        for _ in range(n):
            fn(i)
            i += 1
        e1.record()
        torch.cuda.synchronize()
        spent += e0.elapsed_time(e1)
        if spent >= warm_ms:
            break
        n = min(4096, max(n * 2, 1))
    while True:
        e0.record()
        for i in range(reps):
            fn(i)
        e1.record()
        torch.cuda.synchronize()
        total = e0.elapsed_time(e1)
        if total >= min_ms or reps >= 4096:
            return total / reps
        reps = min(4096, max(reps * 2, int(reps * min_ms / max(total, 1e-3)) + 1))


def _bytes(fi, fo):
    return (8 if fi == "split" else 4) + (8 if fo == "split" else 4)


def run_batched(torch, name, dims, batch, x, fi, fo, scaling, reps, check_transforms, direction=mc.FORWARD,
                vendor=True):
    import oracle
    n = int(np.prod(dims))
    p = mc.Plan(dims, batch, fi, fo, scaling)
    d_in = mc.to_device(x, fi)
    # rotate inputs so the timed region reads HBM, not the Infinity Cache
    rot = max(1, min(4, (2 * MALL) // max(1, d_in.numel() * d_in.element_size()) + 1))
    ins = [d_in] + [d_in.clone() for _ in range(rot - 1)]
    d_out = mc.empty(n * batch, fo)
    # pointers and stream resolved once, as for the rocFFT leg: a few-microsecond
    # exec would otherwise time the Python-side lookups
    st = torch.cuda.current_stream().cuda_stream
    in_ptrs, out_ptr = [t.data_ptr() for t in ins], d_out.data_ptr()
    ms = _timer(torch, lambda i: p.exec_ptr(in_ptrs[i % rot], out_ptr, direction, st), reps)
    flops = 5.0 * n * math.log2(n) * batch
    res = {"config": name, "dims": list(dims), "batch": batch, "in": fi, "out": fo, "scaling": scaling,
           "ms": round(ms, 4), "gflops": round(flops / (ms * 1e-3) / 1e9, 1),
           "hbm_frac": round(n * batch * _bytes(fi, fo) / (ms * 1e-3) / 1e9 / HBM, 4)}
    p.exec(ins[0], d_out, direction)
    torch.cuda.synchronize()
    k = min(batch, check_transforms)
    y = p.result(d_out, k)
    xq = mc.decode(mc.encode(x[: k * n], fi), fi)
    ref = oracle.fft(xq, tuple(dims), sign=-1 if direction == mc.FORWARD else 1)
    res.update({"rel_l2": oracle.rel_l2(y, ref), "max_abs": oracle.max_abs(y, ref), "checked": k})
    p.destroy()
    if vendor:
        res.update(run_rocfft(torch, dims, batch, x, ins, fi, fo, reps, k, ref, direction, ms))
    return res


def run_rocfft(torch, dims, batch, x, ins, fi, fo, reps, k, ref, direction, ours_ms):
    """rocFFT on the same data: half precision (same 4-byte interleaved storage)
    against plain FP16/BF16 runs, single precision against split runs."""
    import oracle
    from mcfft_amd import vendor_baseline as vb
    n = int(np.prod(dims))
    prec = "single" if fo == "split" else "half"
    try:
        rp = vb.RocfftPlan(dims, batch, prec, inverse=(direction == mc.INVERSE))
    except Exception as e:  # noqa: BLE001 -- report, do not fail the run
        return {"rocfft": f"unavailable: {e}"}
    if prec == "half":
        if fi != "fp16":
            return {}
        vins = ins
    else:
        xs = torch.from_numpy(np.ascontiguousarray(mc.decode(mc.encode(x, fi), fi).astype(np.complex64)).view(np.float32))
        vins = [xs.to("cuda").view(-1, 2)]
    vout = torch.empty((n * batch, 2), dtype=vins[0].dtype, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    rot = len(vins)
    vms = _timer(torch, lambda i: rp.exec(vins[i % rot].data_ptr(), vout.data_ptr(), st), reps)
    rp.exec(vins[0].data_ptr(), vout.data_ptr(), st)
    torch.cuda.synchronize()
    yv = vout[: k * n].double().cpu().numpy()
    yv = yv[:, 0] + 1j * yv[:, 1]
    rp.destroy()
    flops = 5.0 * n * math.log2(n) * batch
    return {"rocfft_precision": prec, "rocfft_ms": round(vms, 4), "rocfft_gflops": round(flops / (vms * 1e-3) / 1e9, 1),
            "rocfft_rel_l2": oracle.rel_l2(yv, ref), "speedup_vs_rocfft": round(vms / ours_ms, 3)}


def config_inputs(cfg, rng, dims, batch):
    n = int(np.prod(dims)) * batch
    if cfg in ("c1", "c3", "c6"):
        x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    else:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    if cfg == "c3":  # + amplitude-100 2D exponential at a random (kx, ky) per transform
        nx, ny = dims
        xx = x.reshape(batch, nx, ny)
        m = np.arange(nx)[:, None]
        k = np.arange(ny)[None, :]
        for b in range(batch):
            kx, ky = rng.integers(0, nx), rng.integers(0, ny)
            xx[b] += 100.0 * np.exp(2j * np.pi * (kx * m / nx + ky * k / ny))
        x = xx.reshape(-1)
    return x


def bench_c5(torch, reps, rng, fi="fp16", fo="fp16"):
    """20000 1D transforms, N = 2^k with k uniform in 8..18, plus 64 2D transforms
    with sides from {256, 512, 1024, 2048}; laid out grouped by shape.  Timed as
    phased ragged launches (the default), as one launch with in-kernel pass
    dependencies (MCFFT_RAGGED_SINGLE=1), as per-shape batched launches and as one
    launch per transform, plus the ragged exec replayed from a hipGraph; progress
    goes to stderr."""
    def progress(msg):
        print(f"c5 {fi}->{fo}: {msg}", file=sys.stderr, flush=True)

    ks = rng.integers(8, 19, 20000)
    shapes = [[int(2 ** k)] for k in ks]
    sides = [256, 512, 1024, 2048]
    shapes += [[int(rng.choice(sides)), int(rng.choice(sides))] for _ in range(64)]
    shapes.sort(key=lambda s: (len(s), s))
    sizes = np.array([int(np.prod(s)) for s in shapes], dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sizes.sum())
    flops = float(sum(5.0 * s * math.log2(s) for s in sizes))
    # N(0,1) re/im generated on the device in FP16 (about 1e9 elements)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    d_in = torch.randn((total, 2), dtype=torch.float16, device="cuda", generator=gen)
    d_out = mc.empty(total, fo)
    rp = mc.RaggedPlan([(s, o, o) for s, o in zip(shapes, offs)], fi, fo)
    progress(f"{len(shapes)} items, {total} elements planned")
    ms_r = _timer(torch, lambda i: rp.exec(d_in, d_out, mc.FORWARD), reps)
    flags = rp.status_flags()
    progress(f"phased {ms_r:.3f} ms")
    # the same work as ONE launch with in-kernel pass dependencies
    os.environ["MCFFT_RAGGED_SINGLE"] = "1"
    ms_single = _timer(torch, lambda i: rp.exec(d_in, d_out, mc.FORWARD), reps)
    flags |= rp.status_flags()
    progress(f"single launch {ms_single:.3f} ms")
    del os.environ["MCFFT_RAGGED_SINGLE"]
    # per-shape batched launches over the same buffers
    groups = {}
    for s, o in zip(shapes, offs):
        groups.setdefault(tuple(s), []).append(int(o))
    plans = []
    for s, os_ in groups.items():
        n = int(np.prod(s))
        plans.append((mc.Plan(list(s), len(os_), fi, fo), os_[0], n * len(os_)))
    esz_in, esz_out = d_in.shape[1], d_out.shape[1]

    def per_shape(i):
        for p, o, cnt in plans:
            p.exec_ptr(d_in.data_ptr() + o * esz_in * d_in.element_size(),
                       d_out.data_ptr() + o * esz_out * d_out.element_size(), mc.FORWARD,
                       torch.cuda.current_stream().cuda_stream)

    ms_p = _timer(torch, per_shape, reps)
    progress(f"per-shape {ms_p:.3f} ms")
    # one launch per transform: a batch-1 plan per distinct shape, item by item
    single = {s: mc.Plan(list(s), 1, fi, fo) for s in groups}
    item_ptrs = [(single[tuple(s)], d_in.data_ptr() + int(o) * esz_in * d_in.element_size(),
                  d_out.data_ptr() + int(o) * esz_out * d_out.element_size()) for s, o in zip(shapes, offs)]

    def per_item(i):
        st = torch.cuda.current_stream().cuda_stream
        for p, pi, po in item_ptrs:
            p.exec_ptr(pi, po, mc.FORWARD, st)

    ms_i = _timer(torch, per_item, 2)
    progress(f"per-item {ms_i:.3f} ms")
    # the ragged exec captured in a hipGraph and replayed
    gs = torch.cuda.Stream()
    gs.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(gs):
        with torch.cuda.graph(graph, stream=gs):
            rp.exec(d_in, d_out, mc.FORWARD, stream=gs)
    torch.cuda.synchronize()
    ms_g = _timer(torch, lambda i: graph.replay(), reps)
    progress(f"ragged graph replay {ms_g:.3f} ms")
    # accuracy on a sample of items (ragged output)
    import oracle
    rp.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    sample = rng.choice(len(shapes), 40, replace=False)
    worst = 0.0
    for i in sample:
        s, o, n = shapes[i], int(offs[i]), int(sizes[i])
        y = mc.from_device(d_out[o:o + n], fo, rp.exponents[i])
        xq = mc.from_device(d_in[o:o + n], fi)
        worst = max(worst, oracle.rel_l2(y, oracle.fft(xq, tuple(s))))
    res = {"config": "c5", "items": len(shapes), "elements": total, "in": fi, "out": fo,
           "ragged_ms": round(ms_r, 4), "ragged_gflops": round(flops / (ms_r * 1e-3) / 1e9, 1),
           "ragged_single_launch_ms": round(ms_single, 4),
           "ragged_hbm_frac": round(total * _bytes(fi, fo) / (ms_r * 1e-3) / 1e9 / HBM, 4),
           "per_shape_launches": len(plans), "per_shape_ms": round(ms_p, 4),
           "speedup_vs_per_shape": round(ms_p / ms_r, 3), "ragged_graph_ms": round(ms_g, 4),
           "per_item_launches": len(item_ptrs), "per_item_ms": round(ms_i, 4),
           "speedup_vs_per_item": round(ms_i / ms_r, 3), "status_flags": flags,
           "worst_rel_l2_sample": worst}
    for p, _, _ in plans:
        p.destroy()
    for p in single.values():
        p.destroy()
    rp.destroy()
    return res


def ragged_small(torch, reps, fi="fp16", fo="fp16", seed=1):
    """205 mixed items (1D N = 2^k, k in 4..18, plus five 2D/3D shapes) as one
    phased ragged launch, as per-shape batched launches, as one launch per
    transform, and the first two captured in a hipGraph; the item set of the GPU
    ragged tests."""
    rng = np.random.default_rng(seed)
    shapes = [[int(2 ** k)] for k in rng.integers(4, 19, 200)]
    shapes += [[256, 512], [64, 64], [16, 16, 32], [1024, 256], [32, 2048]]
    shapes.sort(key=lambda s: (len(s), int(np.prod(s))))
    sizes = np.array([int(np.prod(s)) for s in shapes], dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sizes.sum())
    flops = float(sum(5.0 * s * math.log2(s) for s in sizes))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    d_in = torch.randn((total, 2), dtype=torch.float16, device="cuda", generator=gen)
    d_out = mc.empty(total, fo)
    rp = mc.RaggedPlan([(s, o, o) for s, o in zip(shapes, offs)], fi, fo)
    groups = {}
    for s, o in zip(shapes, offs):
        groups.setdefault(tuple(s), []).append(int(o))
    plans = []
    for s, os_ in groups.items():
        n = int(np.prod(s))
        plans.append((mc.Plan(list(s), len(os_), fi, fo), os_[0], n * len(os_)))
    esz_in, esz_out = d_in.shape[1] * d_in.element_size(), d_out.shape[1] * d_out.element_size()

    def per_shape(i):
        st = torch.cuda.current_stream().cuda_stream
        for p, o, cnt in plans:
            p.exec_ptr(d_in.data_ptr() + o * esz_in, d_out.data_ptr() + o * esz_out, mc.FORWARD, st)

    # one launch per transform: a batch-1 plan per distinct shape, executed item by item
    single = {s: mc.Plan(list(s), 1, fi, fo) for s in groups}
    items = [(single[tuple(s)], int(o)) for s, o in zip(shapes, offs)]

    def per_item(i):
        st = torch.cuda.current_stream().cuda_stream
        for p, o in items:
            p.exec_ptr(d_in.data_ptr() + o * esz_in, d_out.data_ptr() + o * esz_out, mc.FORWARD, st)

    ms_r = _timer(torch, lambda i: rp.exec(d_in, d_out, mc.FORWARD), reps)
    ms_p = _timer(torch, per_shape, reps)
    ms_i = _timer(torch, per_item, max(2, reps // 4))
    graphs = {}
    for name, fn in (("ragged", lambda: rp.exec(d_in, d_out, mc.FORWARD)), ("per_shape", lambda: per_shape(0))):
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                fn()
        torch.cuda.synchronize()
        graphs[name] = _timer(torch, lambda i, g=g: g.replay(), reps)
    flags = rp.status_flags()
    res = {"config": "ragged_small", "items": len(shapes), "elements": total, "in": fi, "out": fo,
           "ragged_ms": round(ms_r, 4), "ragged_gflops": round(flops / (ms_r * 1e-3) / 1e9, 1),
           "per_shape_launches": len(plans), "per_shape_ms": round(ms_p, 4),
           "ragged_graph_ms": round(graphs["ragged"], 4), "per_shape_graph_ms": round(graphs["per_shape"], 4),
           "speedup_vs_per_shape": round(ms_p / ms_r, 3),
           "speedup_graph_vs_per_shape_graph": round(graphs["per_shape"] / graphs["ragged"], 3),
           "per_item_launches": len(items), "per_item_ms": round(ms_i, 4),
           "speedup_vs_per_item": round(ms_i / ms_r, 3),
           "status_flags": flags}
    for p, _, _ in plans:
        p.destroy()
    for p in single.values():
        p.destroy()
    rp.destroy()
    return res


def bench_real(torch, reps, n=131072, batch=256):
    """Real-input forward (R2C) of `batch` FP16 real sequences of length n: this
    library (length-n/2 complex FFT + FP32 post-pass) next to rocFFT half
    precision; error against the FP64 oracle on a sample of transforms."""
    import oracle
    from mcfft_amd import vendor_baseline as vb
    rng = np.random.default_rng(7)
    x = rng.standard_normal((batch, n))
    z = mc.real_to_pairs(x)
    d_in = mc.to_device(z, "fp16")
    h = n // 2
    d_out = mc.empty((h + 1) * batch, "fp16")
    p = mc.RealPlan(n, batch, "r2c", "fp16", "fp16")
    ms = _timer(torch, lambda i: p.exec(d_in, d_out, mc.FORWARD), reps)
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    k = 4
    y = p.result(d_out, k)
    xq = mc.pairs_to_real(mc.decode(mc.encode(z[: k * h], "fp16"), "fp16"), n)
    ref = oracle.fft(xq.reshape(-1).astype(np.complex128), (n,)).reshape(k, n)[:, : h + 1].reshape(-1)
    flops = 2.5 * n * math.log2(n) * batch   # half of the complex transform's 5 N log2 N
    res = {"config": "r2c", "n": n, "batch": batch, "in": "fp16", "out": "fp16", "ms": round(ms, 4),
           "gflops_2.5nlogn": round(flops / (ms * 1e-3) / 1e9, 1),
           "hbm_frac": round((2 * n + 4 * (h + 1)) * batch / (ms * 1e-3) / 1e9 / HBM, 4),
           "rel_l2": oracle.rel_l2(y, ref), "checked": k}
    p.destroy()
    try:
        rp = vb.RocfftPlan([n], batch, "half", real=True)
        xr = torch.from_numpy(x.astype(np.float16)).to("cuda")
        vout = torch.empty(((h + 1) * batch, 2), dtype=torch.float16, device="cuda")
        st = torch.cuda.current_stream().cuda_stream
        vms = _timer(torch, lambda i: rp.exec(xr.data_ptr(), vout.data_ptr(), st), reps)
        res.update({"rocfft_half_ms": round(vms, 4), "speedup_vs_rocfft": round(vms / ms, 3)})
        rp.destroy()
    except Exception as e:  # noqa: BLE001
        res["rocfft"] = f"unavailable: {e}"
    return res


def bench_layout(torch, reps, n=131072, batch=256):
    """Planar FP16 input and output (separate re / im arrays) of the C2 shape:
    gather + transform + scatter next to the contiguous interleaved exec, and
    the HBM rate of the staging kernels (4 B read + 4 B written per element
    each)."""
    g = torch.Generator(device="cuda").manual_seed(3)
    inter = torch.randn((n * batch, 2), device="cuda", generator=g).half()
    re, im = inter[:, 0].contiguous(), inter[:, 1].contiguous()
    o_re, o_im = torch.empty_like(re), torch.empty_like(im)
    out = torch.empty_like(inter)
    lay = mc.Layout.make([1], n, True)
    p = mc.LayoutPlan([n], batch, lay, lay, "fp16", "fp16")
    q = mc.Plan([n], batch, "fp16", "fp16")
    ms_l = _timer(torch, lambda i: p.exec_layout(re, im, o_re, o_im, mc.FORWARD), reps)
    ms_c = _timer(torch, lambda i: q.exec(inter, out, mc.FORWARD), reps)
    q.exec(inter, out, mc.FORWARD)
    p.exec_layout(re, im, o_re, o_im, mc.FORWARD)
    torch.cuda.synchronize()
    same = bool(torch.equal(o_re.view(torch.int16), out[:, 0].contiguous().view(torch.int16)) and
                torch.equal(o_im.view(torch.int16), out[:, 1].contiguous().view(torch.int16)))
    stage_ms = ms_l - ms_c
    res = {"config": "layout_planar", "n": n, "batch": batch, "fmt": "fp16", "ms_layout": round(ms_l, 4),
           "ms_contiguous": round(ms_c, 4), "staging_ms": round(stage_ms, 4),
           "staging_gbs": round(2 * 8 * n * batch / (stage_ms * 1e-3) / 1e9, 1) if stage_ms > 0 else None,
           "bit_identical": same}
    p.destroy()
    # spaced interleaved transforms (64 elements between them): no staging pass
    dist = n + 64
    sp_in = torch.zeros(((batch - 1) * dist + n, 2), device="cuda", dtype=torch.float16)
    for t in range(batch):
        sp_in[t * dist:t * dist + n] = inter[t * n:(t + 1) * n]
    sp_out = torch.empty_like(sp_in)
    lay = mc.Layout.make([1], dist)
    ps = mc.LayoutPlan([n], batch, lay, lay, "fp16", "fp16")
    ms_s = _timer(torch, lambda i: ps.exec_layout(sp_in, None, sp_out, None, mc.FORWARD), reps)
    ps.exec_layout(sp_in, None, sp_out, None, mc.FORWARD)
    torch.cuda.synchronize()
    same_s = all(bool(torch.equal(sp_out[t * dist:t * dist + n].view(torch.int32), out[t * n:(t + 1) * n].view(torch.int32)))
                 for t in range(batch))
    res["spaced"] = {"dist": dist, "ms": round(ms_s, 4), "staging_bytes": ps.workspace_bytes() - q.workspace_bytes(),
                     "bit_identical": same_s}
    ps.destroy()
    q.destroy()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c1,c2,c3,c4,c5,c6,ragged_small,real,layout")
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()
    import torch
    cfgs = args.config.split(",")
    rng = np.random.default_rng(1)
    table = {
        "c1": ([256], 4096, "auto", 4096),
        "c2": ([131072], 256, "auto", 8),
        "c3": ([4096, 4096], 4, "auto", 1),
        "c4": ([512, 512, 512], 1, "auto", 1),
        "c6": ([65536], 4096, "auto", 16),
    }
    for cfg in cfgs:
        if cfg == "layout":
            print(json.dumps(bench_layout(torch, args.reps)), flush=True)
            continue
        if cfg == "real":
            print(json.dumps(bench_real(torch, args.reps)), flush=True)
            continue
        if cfg == "ragged_small":
            for fi, fo in (("fp16", "fp16"), ("fp16", "split")):
                print(json.dumps(ragged_small(torch, args.reps, fi, fo)), flush=True)
            continue
        if cfg == "c5":
            for fi, fo in (("fp16", "fp16"), ("fp16", "split")):
                print(json.dumps(bench_c5(torch, args.reps, np.random.default_rng(5), fi, fo)), flush=True)
            continue
        dims, batch, scaling, check = table[cfg]
        x = config_inputs(cfg, rng, dims, batch)
        precs = [("fp16", "fp16"), ("fp16", "split")]
        if cfg == "c3":
            precs.append(("bf16", "bf16"))
        for fi, fo in precs:
            r = run_batched(torch, cfg, dims, batch, x, fi, fo, scaling, args.reps, check)
            print(json.dumps(r), flush=True)
        if cfg == "c2":  # the inverse leg of the forward+inverse config
            r = run_batched(torch, "c2-inverse", dims, batch, x, "split", "split", "auto", args.reps, check,
                            direction=mc.INVERSE)
            print(json.dumps(r), flush=True)
        del x


if __name__ == "__main__":
    main()
