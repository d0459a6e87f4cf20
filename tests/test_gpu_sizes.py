"""GPU: every supported 1D length (16 .. 2^27) in each format, and the NONE and
FULL scaling modes (plans and ragged batches), against the FP64 oracle; a
forward + inverse round trip at 2^27 checked in FP64 on the GPU.

Tolerances as in the parity tests: rel-L2 fp16 1.5e-3, bf16 1.2e-2, split 1e-6.
Scaling: NONE stores the unnormalised result (exponent 0), FULL stores it
divided by N (exponent log2 N)."""
import math

import numpy as np
import pytest

import mcfft_amd as mc
import oracle

pytestmark = pytest.mark.gpu

TOL = {"fp16": 1.5e-3, "bf16": 1.2e-2, "split": 1e-6}


def _fmts(fmt):
    return ("fp16", "split") if fmt == "split" else (fmt, fmt)


_REF = {}


def _reference(x, in_fmt, dims, key):
    """FP64 oracle FFT of the quantised input; the last two kept (the 2^24..2^27
    cases of FP16 and split share one)."""
    if key not in _REF:
        if len(_REF) >= 2:
            _REF.pop(next(iter(_REF)))
        _REF[key] = oracle.fft(mc.decode(mc.encode(x, in_fmt), in_fmt), tuple(dims))
    return _REF[key]


def _run(torch, dims, batch, fmt, scaling, seed):
    in_fmt, out_fmt = _fmts(fmt)
    n = int(np.prod(dims))
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)
    p = mc.Plan(list(dims), batch, in_fmt, out_fmt, scaling)
    d_in = mc.to_device(x, in_fmt)
    d_out = mc.empty(n * batch, out_fmt)
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    e = p.exponent
    y = p.result(d_out)
    p.destroy()
    del d_in, d_out
    ref = _reference(x, in_fmt, dims, (tuple(dims), batch, "bf16" if in_fmt == "bf16" else "fp16", seed))
    return oracle.rel_l2(y, ref), e


@pytest.mark.parametrize("lg", list(range(4, 24)))
@pytest.mark.parametrize("fmt", ["fp16", "bf16", "split"])
@pytest.mark.parametrize("scaling", ["auto", "balanced"])
def test_every_1d_length(torch_cuda, lg, fmt, scaling):
    """Each power of two the 1D planner accepts; at least 2^18 elements per case."""
    n = 1 << lg
    batch = max(1, (1 << 18) // n)
    print(f"1d n=2^{lg} b={batch} {fmt} {scaling}", flush=True)
    err, _ = _run(torch_cuda, [n], batch, fmt, scaling, seed=lg)
    print(f"  rel_l2 {err:.3e}", flush=True)
    assert err <= TOL[fmt]


@pytest.mark.parametrize("lg", [24, 25, 26, 27])
@pytest.mark.parametrize("fmt", ["fp16", "split", "bf16"])
def test_every_1d_length_three_pass(torch_cuda, lg, fmt):
    """2^24 .. 2^27 (three-pass plans), batch 1, default scaling; 2^24 also
    with static BALANCED scaling."""
    n = 1 << lg
    for scaling in (["auto", "balanced"] if lg == 24 else ["auto"]):
        print(f"1d n=2^{lg} b=1 {fmt} {scaling}", flush=True)
        err, _ = _run(torch_cuda, [n], 1, fmt, scaling, seed=lg)
        print(f"  rel_l2 {err:.3e}", flush=True)
        assert err <= TOL[fmt]


@pytest.mark.parametrize("fmt", ["fp16", "split"])
def test_roundtrip_2_27(torch_cuda, fmt):
    """N = 2^27 x 2: forward then inverse (the inverse takes the forward's
    exponents), inv(fwd(x)) = N x checked per transform in FP64 on the GPU;
    the second transform runs in place."""
    torch = torch_cuda
    n, batch = 1 << 27, 2
    in_fmt, out_fmt = _fmts(fmt)
    g = torch.Generator(device="cuda")
    g.manual_seed(27)
    x = torch.randn((n * batch, 2), dtype=torch.float16, device="cuda", generator=g)
    ref = x.double().reshape(-1)
    pf = mc.Plan([n], batch, in_fmt, out_fmt)
    pi = mc.Plan([n], batch, out_fmt, out_fmt)
    mid, out = mc.empty(n * batch, out_fmt), mc.empty(n * batch, out_fmt)
    pf.exec(x, mid, mc.FORWARD)
    pi.exec(mid, out, mc.INVERSE, in_exp=pf)
    e = pi.exponents()
    assert pf.status_flags() == 0 and pi.status_flags() == 0
    for b in range(batch):
        rel, _ = mc.error_norms(out[b * n:(b + 1) * n], out_fmt, int(e[b]) - 27, ref[2 * b * n:2 * (b + 1) * n])
        print(f"2^27 round trip {fmt} transform {b}: rel_l2 {rel:.3e}", flush=True)
        assert rel <= (2e-6 if fmt == "split" else 3e-3)
    # in place (input = output buffer) gives the same values
    pi.exec(mid, mid, mc.INVERSE, in_exp=pf)
    torch.cuda.synchronize()
    assert torch.equal(mid.view(torch.int16), out.view(torch.int16))
    pf.destroy()
    pi.destroy()


@pytest.mark.parametrize("dims,batch", [((256,), 64), ((4096,), 8), ((131072,), 2), ((1 << 20,), 1),
                                        ((256, 256), 2), ((4096, 4096), 1), ((64, 64, 64), 1)])
@pytest.mark.parametrize("fmt", ["fp16", "bf16", "split"])
def test_scaling_full(torch_cuda, dims, batch, fmt):
    n = int(np.prod(dims))
    print(f"full scaling dims={dims} b={batch} {fmt}", flush=True)
    err, e = _run(torch_cuda, dims, batch, fmt, "full", seed=n + 1)
    print(f"  rel_l2 {err:.3e} exponent {e}", flush=True)
    assert e == int(math.log2(n))
    assert err <= TOL[fmt]


@pytest.mark.parametrize("dims,batch", [((256,), 64), ((4096,), 8), ((65536,), 2), ((64, 64), 4),
                                        ((16, 16, 32), 4)])
@pytest.mark.parametrize("fmt", ["fp16", "bf16", "split"])
def test_scaling_none(torch_cuda, dims, batch, fmt):
    """Unnormalised output; N(0,1) input at these sizes stays inside FP16's range."""
    n = int(np.prod(dims))
    print(f"no scaling dims={dims} b={batch} {fmt}", flush=True)
    err, e = _run(torch_cuda, dims, batch, fmt, "none", seed=n + 2)
    print(f"  rel_l2 {err:.3e}", flush=True)
    assert e == 0
    assert err <= TOL[fmt]


@pytest.mark.parametrize("scaling", ["none", "full"])
def test_ragged_scaling_modes(torch_cuda, scaling):
    """Per-item exponents of a ragged batch follow the scaling mode (0 / log2 N)."""
    torch = torch_cuda
    shapes = [[16], [4096], [256, 512], [131072], [64, 64], [1024], [16, 16, 32], [32768]]
    sizes = [int(np.prod(s)) for s in shapes]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sum(sizes))
    print(f"ragged scaling={scaling}: {len(shapes)} items", flush=True)
    rng = np.random.default_rng(21)
    x = rng.standard_normal(total) + 1j * rng.standard_normal(total)
    p = mc.RaggedPlan([(s, int(o), int(o)) for s, o in zip(shapes, offs)], "fp16", "split", scaling)
    d_in = mc.to_device(x, "fp16")
    d_out = mc.empty(total, "split")
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    assert p.status_flags() == 0
    y = mc.from_device(d_out, "split")
    xq = mc.decode(mc.encode(x, "fp16"), "fp16")
    for s, n, o, e in zip(shapes, sizes, offs, p.exponents):
        assert e == (0 if scaling == "none" else int(math.log2(n))), (s, e)
        err = oracle.rel_l2(y[o:o + n] * 2.0 ** e, oracle.fft(xq[o:o + n], tuple(s)))
        assert err <= TOL["split"], (s, err)
    p.destroy()


@pytest.mark.parametrize("dims,batch,fmt", [((256, 16), 8, "fp16"), ((256, 16), 8, "bf16"), ((256, 32), 4, "fp16"),
                                            ((256, 32), 4, "bf16"), ((256, 32), 4, "split"),
                                            ((256, 256, 16), 1, "fp16"), ((256, 256, 16), 1, "bf16"),
                                            ((256, 16), 8, "split"), ((256, 256, 16), 1, "split"),
                                            ((512, 16), 4, "fp16"), ((512, 16), 4, "bf16"),
                                            ((32, 16), 8, "fp16"), ((64, 16), 8, "bf16"), ((64, 16), 8, "split"),
                                            ((128, 16), 4, "fp16"), ((64, 64, 16), 2, "fp16"),
                                            ((128, 64, 16), 1, "split"), ((2048, 32), 2, "split")])
def test_narrow_outer_axes(torch_cuda, dims, batch, fmt):
    """Outer axes with fewer columns than the preferred column tile (32 or 64)
    run on a compiled variant with a 16-column tile."""
    print(f"narrow dims={dims} b={batch} {fmt}", flush=True)
    err, _ = _run(torch_cuda, dims, batch, fmt, "balanced", seed=int(np.prod(dims)) + 3)
    print(f"  rel_l2 {err:.3e}", flush=True)
    assert err <= TOL[fmt]
