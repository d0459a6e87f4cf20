"""GPU parity: libmcfft.so (through its C ABI) against the FP64 oracle.

Tolerances (rel-L2 against an FP64 FFT of the same quantised input):
  fp16 plain   <= 1.5e-3 a-priori bound (~2^-11 per stage), and no worse than the
               reference algorithm emulated at FP16 where the reference has a plan
  bf16 plain   <= 1.2e-2 (~2^-8 per stage)
  fp16 split   <= 1e-6
"""
import numpy as np
import pytest

import mcfft_amd as mc
import oracle

pytestmark = pytest.mark.gpu

TOL = {"fp16": 1.5e-3, "bf16": 1.2e-2, "split": 1e-6}


def _input(n_total, seed, kind="normal"):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        return rng.uniform(-1, 1, n_total) + 1j * rng.uniform(-1, 1, n_total)
    return rng.standard_normal(n_total) + 1j * rng.standard_normal(n_total)


def _quantised(x, in_fmt):
    return mc.decode(mc.encode(x, in_fmt), in_fmt)


def run(dims, batch, in_fmt, out_fmt, x, direction=mc.FORWARD, scaling="auto", inplace=False):
    import torch
    p = mc.Plan(dims, batch, in_fmt, out_fmt, scaling)
    din = mc.to_device(x, in_fmt)
    if inplace:
        assert in_fmt == out_fmt
        dout = din
    else:
        dout = mc.empty(x.size, out_fmt)
    p.exec(din, dout, direction)
    torch.cuda.synchronize()
    y = p.result(dout)
    p.destroy()
    return y


CASES_1D = [(16, 64), (32, 33), (64, 16), (128, 8), (256, 4096), (512, 16), (1024, 8), (2048, 4), (4096, 3),
            (8192, 2), (65536, 2), (131072, 2), (1 << 20, 1)]


@pytest.mark.parametrize("n,batch", CASES_1D)
@pytest.mark.parametrize("fmt", ["fp16", "bf16", "split"])
def test_1d_forward(torch_cuda, n, batch, fmt):
    in_fmt, out_fmt = ("fp16", "split") if fmt == "split" else (fmt, fmt)
    x = _input(n * batch, seed=n + batch)
    xq = _quantised(x, in_fmt)
    y = run([n], batch, in_fmt, out_fmt, x)
    ref = oracle.fft(xq, (n,))
    err = oracle.rel_l2(y, ref)
    print(f"1d n={n} b={batch} {fmt} rel_l2={err:.3e}")
    assert err <= TOL[fmt]


@pytest.mark.parametrize("dims,batch", [((64, 32), 3), ((256, 256), 2), ((1024, 512), 1), ((4096, 4096), 1),
                                        ((512, 256), 2), ((16, 65536), 1), ((16, 16, 32), 2), ((64, 32, 128), 1),
                                        ((128, 128, 128), 1), ((512, 16, 64), 1), ((32, 131072), 2),
                                        ((16, 16, 32768), 1), ((64, 32768), 3)])
@pytest.mark.parametrize("fmt", ["fp16", "bf16", "split"])
def test_nd_forward(torch_cuda, dims, batch, fmt):
    """2D/3D forward against the oracle; innermost sides above the single-pass
    limit (split 2^15 and 2^16, FP16/BF16 2^17) run a two-pass four-step over
    every row before the column passes."""
    in_fmt, out_fmt = ("fp16", "split") if fmt == "split" else (fmt, fmt)
    n = int(np.prod(dims))
    x = _input(n * batch, seed=n)
    xq = _quantised(x, in_fmt)
    y = run(list(dims), batch, in_fmt, out_fmt, x)
    ref = oracle.fft(xq, dims)
    err = oracle.rel_l2(y, ref)
    print(f"nd dims={dims} b={batch} {fmt} rel_l2={err:.3e}")
    assert err <= TOL[fmt]


@pytest.mark.parametrize("dims,batch", [((256,), 8), ((4096,), 2), ((131072,), 1), ((64, 128), 2), ((32, 32, 32), 1)])
def test_inverse_split(torch_cuda, dims, batch):
    n = int(np.prod(dims))
    x = _input(n * batch, seed=7)
    xq = _quantised(x, "fp16")
    y = run(list(dims), batch, "fp16", "split", x, direction=mc.INVERSE)
    ref = oracle.fft(xq, dims, sign=+1)
    assert oracle.rel_l2(y, ref) <= TOL["split"]


def test_inplace_matches_out_of_place(torch_cuda):
    x = _input(65536 * 4, seed=3)
    a = run([65536], 4, "fp16", "fp16", x)
    b = run([65536], 4, "fp16", "fp16", x, inplace=True)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [256, 512, 1024, 131072])
def test_plain_no_worse_than_reference_fp16(torch_cuda, n):
    """One-sided bound: plain FP16 is at least as accurate as the reference's
    algorithm run at MCTYPE=half (emulated by the oracle), on the same inputs.
    What rocWMMA does with a half accumulator is bracketed by two emulations
    (each mma_sync rounds to half; a complex_mul's four mma_sync share an FP32
    accumulator and round once): held to the tighter of the two."""
    batch = max(1, 4096 // n)
    x = _input(n * batch, seed=11)
    xq = _quantised(x, "fp16")
    ref = oracle.fft(xq, (n,))
    y = run([n], batch, "fp16", "fp16", x)
    ours = oracle.rel_l2(y, ref)
    per_mma = oracle.rel_l2(oracle.ref_fp16_exec(xq, n), ref)
    fused = oracle.rel_l2(oracle.ref_fp16_exec(xq, n, points=oracle.R_ALL | oracle.R_FUSED), ref)
    print(f"n={n} ours={ours:.3e} reference-fp16 per-mma={per_mma:.3e} fused={fused:.3e}")
    assert ours <= min(per_mma, fused)


@pytest.mark.parametrize("scaling", ["auto", "balanced"])
def test_split_roundtrip_c2_full_size(torch_cuda, scaling):
    """1D N=131072 x 256 forward then inverse (split): inv(fwd(x)) = N x.  The
    inverse takes the forward's per-transform exponents as its input exponents,
    so its own exponents give the true result."""
    import torch
    n, b = 131072, 256
    x = _input(n * b, seed=2)
    xq = _quantised(x, "fp16")
    pf = mc.Plan([n], b, "fp16", "split", scaling)
    pi = mc.Plan([n], b, "split", "split", scaling)
    d0 = mc.to_device(x, "fp16")
    d1 = mc.empty(n * b, "split")
    d2 = mc.empty(n * b, "split")
    pf.exec(d0, d1, mc.FORWARD)
    pi.exec(d1, d2, mc.INVERSE, in_exp=pf)
    torch.cuda.synchronize()
    y = pi.result(d2) / n
    err = oracle.rel_l2(y, xq)
    print(f"roundtrip {scaling} rel_l2={err:.3e}")
    assert err <= 2e-6


@pytest.mark.parametrize("fmt", ["fp16", "bf16", "split"])
def test_gpu_error_norms_match_host(torch_cuda, fmt):
    """The FP64 GPU error-norm reduction agrees with the host oracle's norms."""
    torch = torch_cuda
    rng = np.random.default_rng(21)
    n = 300001
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = x + 1e-3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    d = mc.to_device(x, fmt)
    dref = torch.from_numpy(ref.view(np.float64).copy()).cuda()
    r, m = mc.error_norms(d, fmt, 3, dref)
    y = mc.decode(mc.encode(x, fmt), fmt, 3)
    assert abs(r - oracle.rel_l2(y, ref)) <= 1e-12 * max(1.0, r)
    assert abs(m - oracle.max_abs(y, ref)) <= 1e-12 * max(1.0, m)


@pytest.mark.parametrize("kernel,fmt,scaling", [("mcfft_pass_big_0", "fp16", "balanced"),
                                                ("mcfft_pass_big16_0", "fp16", "balanced"),
                                                ("mcfft_pass_big16_1", "bf16", "balanced"),
                                                ("mcfft_pass_bigauto_0", "fp16", "auto"),
                                                ("mcfft_pass_bigauto16_0", "fp16", "auto"),
                                                ("mcfft_pass_bigauto_1", "bf16", "auto")])
def test_single_pass_65536_variants(torch_cuda, kernel, fmt, scaling, monkeypatch):
    """Both wave counts of the single-pass 65536 kernel, statically scaled and
    AUTO forms (pinned with MCFFT_FORCE_KERNEL), against the oracle."""
    monkeypatch.setenv("MCFFT_FORCE_KERNEL", kernel)
    n, batch = 65536, 3
    x = _input(n * batch, seed=11)
    p = mc.Plan([n], batch, fmt, fmt, scaling)
    assert kernel in p.describe()
    p.destroy()
    y = run([n], batch, fmt, fmt, x, scaling=scaling)
    err = oracle.rel_l2(y, oracle.fft(_quantised(x, fmt), (n,)))
    print(f"65536 {kernel} rel_l2={err:.3e}")
    assert err <= TOL[fmt]


@pytest.mark.parametrize("scaling", ["auto", "balanced"])
def test_single_pass_65536_stagger_keeps_bits(torch_cuda, monkeypatch, scaling):
    """The single-pass 65536 kernel's start stagger (half the workgroups start
    late when each has four or more transforms: 1024 here) changes timing only:
    output and exponents bit-identical with it off, and one transform against
    the oracle."""
    torch = torch_cuda
    n, batch = 65536, 1024
    g = torch.Generator(device="cuda")
    g.manual_seed(17)
    d_in = torch.randn((n * batch, 2), device="cuda", generator=g).half()
    outs, exps = [], []
    for stagger in ("0", None):
        if stagger is None:
            monkeypatch.delenv("MCFFT_BIG_STAGGER", raising=False)
        else:
            monkeypatch.setenv("MCFFT_BIG_STAGGER", stagger)
        p = mc.Plan([n], batch, "fp16", "fp16", scaling)
        assert "mcfft_pass_big" in p.describe()
        d_out = torch.empty_like(d_in)
        p.exec(d_in, d_out, mc.FORWARD)
        torch.cuda.synchronize()
        outs.append(d_out)
        exps.append(p.exponents().tolist() if scaling == "auto" else [p.exponent])
        p.destroy()
    assert torch.equal(outs[0].view(torch.int16), outs[1].view(torch.int16))
    assert exps[0] == exps[1]
    t = 517
    e = exps[1][t] if scaling == "auto" else exps[1][0]
    y = mc.from_device(outs[1][t * n:(t + 1) * n], "fp16") * 2.0 ** e
    ref = oracle.fft(mc.from_device(d_in[t * n:(t + 1) * n], "fp16"), (n,))
    err = oracle.rel_l2(y, ref)
    print(f"65536 x {batch} {scaling} stagger: bit-identical, transform {t} rel_l2 {err:.3e}")
    assert err <= TOL["fp16"]
