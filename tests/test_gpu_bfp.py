"""GPU: block floating point (the default MCFFT_SCALE_AUTO) on the inputs that
static scaling cannot serve with one setting, through the C ABI, against the
FP64 oracle.

  * The 2D benchmark input: 4096 x 4096, re/im iid U[-1,1) plus
    100 * exp(2 pi i (kx m + ky n) / 4096) at a random (kx, ky) per transform
    (seed 3).  Its spectrum peaks at 100 * 4096^2 ~ 1.7e9, so BALANCED's 2^-12
    overflows FP16 (the status word says so), while FULL's 2^-24 would lose the
    small inputs below.
  * The reference's generate_data input (source/mcfft_1d_utils.cpp:88-100):
    re/im 1e-4 * U[0, 1), partly FP16 subnormals.
Tolerances (rel-L2 against an FP64 FFT of the quantised input): fp16 1.5e-3,
bf16 1.2e-2, split 1e-6 forward; round trips fp16 3e-3, bf16 2.4e-2, split 2e-6.
"""
import math

import numpy as np
import pytest

import mcfft_amd as mc
import oracle

pytestmark = pytest.mark.gpu

TOL = {"fp16": 1.5e-3, "bf16": 1.2e-2, "split": 1e-6}
TOL_RT = {"fp16": 3e-3, "bf16": 2.4e-2, "split": 2e-6}
C3 = (4096, 4096)


def c3_input(batch, seed=3):
    """The 2D benchmark input (see the module docstring)."""
    rng = np.random.default_rng(seed)
    nx, ny = C3
    n = nx * ny * batch
    x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).reshape(batch, nx, ny)
    m = np.arange(nx)[:, None]
    k = np.arange(ny)[None, :]
    for b in range(batch):
        kx, ky = rng.integers(0, nx), rng.integers(0, ny)
        x[b] += 100.0 * np.exp(2j * np.pi * (kx * m / nx + ky * k / ny))
    return x.reshape(-1)


def generate_data_input(n_total, seed=42):
    """re/im 1e-4 * U[0, 1): generate_data's magnitude."""
    rng = np.random.default_rng(seed)
    return 1e-4 * (rng.uniform(0, 1, n_total) + 1j * rng.uniform(0, 1, n_total))


def _fmts(fmt):
    return ("fp16", "split") if fmt == "split" else (fmt, fmt)


def forward(torch, x, dims, batch, fmt, scaling="auto"):
    fi, fo = _fmts(fmt)
    p = mc.Plan(list(dims), batch, fi, fo, scaling)
    d_in = mc.to_device(x, fi)
    d_out = mc.empty(x.size, fo)
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    y = p.result(d_out)
    flags = p.status_flags()
    e = p.exponents()
    p.destroy()
    return y, flags, e, mc.decode(mc.encode(x, fi), fi)


@pytest.mark.parametrize("fmt", ["fp16", "bf16", "split"])
def test_c3_input_forward_default_scaling(torch_cuda, fmt):
    """4096^2 x 1 with the amplitude-100 sinusoid, default plan, vs the oracle."""
    x = c3_input(1)
    y, flags, e, xq = forward(torch_cuda, x, C3, 1, fmt)
    err = oracle.rel_l2(y, oracle.fft(xq, C3))
    print(f"c3 {fmt}: rel_l2 {err:.3e} exponent {e[0]} flags {flags}")
    assert flags == 0
    assert err <= TOL[fmt]


@pytest.mark.parametrize("fmt", ["fp16", "bf16", "split"])
def test_c3_input_roundtrip_batch4_default_scaling(torch_cuda, fmt):
    """4096^2 x 4 forward then inverse (the inverse takes the forward's
    per-transform exponents as its input exponents), each transform checked in
    FP64 on the GPU: inv(fwd(x)) = N x."""
    torch = torch_cuda
    batch, n = 4, C3[0] * C3[1]
    x = c3_input(batch)
    fi, fo = _fmts(fmt)
    pf = mc.Plan(list(C3), batch, fi, fo)
    pi = mc.Plan(list(C3), batch, fo, fo)
    d_in = mc.to_device(x, fi)
    ref = torch.from_numpy(mc.decode(mc.encode(x, fi), fi).view(np.float64).copy()).cuda()
    del x
    d_mid, d_out = mc.empty(n * batch, fo), mc.empty(n * batch, fo)
    pf.exec(d_in, d_mid, mc.FORWARD)
    pi.exec(d_mid, d_out, mc.INVERSE, in_exp=pf)
    e = pi.exponents()
    assert pf.status_flags() == 0 and pi.status_flags() == 0
    for b in range(batch):
        rel, _ = mc.error_norms(d_out[b * n:(b + 1) * n], fo, int(e[b]) - int(math.log2(n)), ref[2 * b * n:2 * (b + 1) * n])
        print(f"c3 x4 round trip {fmt} transform {b}: rel_l2 {rel:.3e} exponent {e[b]}")
        assert rel <= TOL_RT[fmt]
    pf.destroy()
    pi.destroy()


@pytest.mark.parametrize("fmt", ["fp16", "split"])
def test_c3_input_balanced_reports_overflow(torch_cuda, fmt):
    """Static BALANCED scaling overflows FP16 on this input: the stored result
    holds Inf and the status word reports it."""
    x = c3_input(1)
    y, flags, _, _ = forward(torch_cuda, x, C3, 1, fmt, scaling="balanced")
    assert flags & mc.STATUS_NONFINITE
    assert not np.isfinite(y).all()


@pytest.mark.parametrize("dims,batch", [(C3, 1), ((256,), 4096), ((131072,), 4)])
@pytest.mark.parametrize("fmt", ["fp16", "split"])
def test_generate_data_input_default_scaling(torch_cuda, dims, batch, fmt):
    """generate_data-sized inputs (~1e-4): the default plan scales them up."""
    n = int(np.prod(dims))
    x = generate_data_input(n * batch)
    y, flags, e, xq = forward(torch_cuda, x, dims, batch, fmt)
    err = oracle.rel_l2(y, oracle.fft(xq, dims))
    print(f"generate_data dims={dims} {fmt}: rel_l2 {err:.3e} exponents {e.min()}..{e.max()}")
    assert flags == 0
    assert err <= TOL[fmt]


@pytest.mark.parametrize("case", ["2d", "3d", "fourstep"])
@pytest.mark.parametrize("fmt", ["fp16", "split"])
def test_auto_reconcile(torch_cuda, case, fmt):
    """Tiles of the first pass whose input maxima lie 2^4..2^12 apart choose
    different exponents; the next pass brings them to one (vs the oracle)."""
    rng = np.random.default_rng(17)
    if case == "2d":
        dims, batch = (512, 1024), 2
        scale = np.exp2(4.0 * (np.arange(512) % 4))[:, None] * np.ones((1, 1024))
    elif case == "3d":
        dims, batch = (32, 64, 128), 2
        scale = np.exp2(4.0 * (np.arange(64) % 3))[None, :, None] * np.ones((32, 64, 128))
    else:
        dims, batch = (131072,), 4
        scale = np.exp2(4.0 * ((np.arange(131072) % 512) // 32 % 3))
    n = int(np.prod(dims))
    x = (rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)) * np.tile(scale.reshape(-1), batch)
    y, flags, e, xq = forward(torch_cuda, x, dims, batch, fmt)
    err = oracle.rel_l2(y, oracle.fft(xq, dims))
    print(f"reconcile {case} {fmt}: rel_l2 {err:.3e}")
    assert flags == 0
    assert err <= TOL[fmt]


@pytest.mark.parametrize("dims,batch", [((256,), 8), ((65536,), 4), ((131072,), 4), ((64, 256), 4)])
def test_auto_nonfinite_and_zero_transforms(torch_cuda, dims, batch):
    """An Inf in transform 1's input sets the status bit (read once, then
    cleared); an all-zero transform 2 comes out as zeros; transforms 0 and 3
    keep their tolerance."""
    n = int(np.prod(dims))
    rng = np.random.default_rng(8)
    x = (rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)).reshape(batch, n)
    x[1, 7] = np.inf
    x[2] = 0
    y, flags, e, xq = forward(torch_cuda, x.reshape(-1), dims, batch, "fp16")
    assert flags & mc.STATUS_NONFINITE
    y, xq = y.reshape(batch, n), xq.reshape(batch, n)
    assert (y[2] == 0).all()
    for b in (0, 3):
        assert oracle.rel_l2(y[b], oracle.fft(xq[b], dims)) <= TOL["fp16"]


def test_status_is_cleared_by_reading(torch_cuda):
    torch = torch_cuda
    p = mc.Plan([1024], 2, "fp16", "fp16", "none")
    x = np.full(2048, 1000.0 + 0j)   # DC bin 1.0e6: overflows FP16 unscaled
    d_in, d_out = mc.to_device(x, "fp16"), mc.empty(2048, "fp16")
    p.exec(d_in, d_out, mc.FORWARD)
    assert p.status_flags() & mc.STATUS_NONFINITE
    assert p.status_flags() == 0
    p.exec(mc.to_device(np.ones(2048, complex), "fp16"), d_out, mc.FORWARD)
    torch.cuda.synchronize()
    assert p.status_flags() == 0
    p.destroy()


@pytest.mark.parametrize("scaling", ["auto", "balanced"])
def test_input_exponents(torch_cuda, scaling):
    """Per-transform input exponents add to the output exponents (and nothing
    else changes), through mcfft_exec_exp."""
    torch = torch_cuda
    n, batch = 131072, 4
    rng = np.random.default_rng(6)
    x = rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)
    p = mc.Plan([n], batch, "fp16", "fp16", scaling)
    d_in, d0, d1 = mc.to_device(x, "fp16"), mc.empty(n * batch, "fp16"), mc.empty(n * batch, "fp16")
    p.exec(d_in, d0, mc.FORWARD)
    e0 = p.exponents()
    ins = torch.tensor([3, -2, 0, 9], dtype=torch.int32, device="cuda")
    p.exec(d_in, d1, mc.FORWARD, in_exp=ins)
    e1 = p.exponents()
    assert (e1 == e0 + np.array([3, -2, 0, 9])).all()
    assert torch.equal(d0, d1)
    p.destroy()
