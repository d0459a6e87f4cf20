"""GPU: real-input forward (R2C) and real-output inverse (C2R) transforms, 1D,
2D and 3D, against the FP64 oracle (the complex FFT of the real input / of the
Hermitian-extended spectrum)."""
import numpy as np
import pytest

import mcfft_amd as mc
import oracle

pytestmark = pytest.mark.gpu
TOL = {"fp16": 1.5e-3, "bf16": 1.2e-2, "split": 1e-6}
FMTS = [("fp16", "fp16"), ("bf16", "bf16"), ("fp16", "split")]


@pytest.mark.parametrize("n,batch", [(32, 64), (256, 512), (4096, 16), (65536, 4), (131072, 2), (1 << 20, 1)])
@pytest.mark.parametrize("fi,fo", FMTS)
def test_r2c(torch_cuda, n, batch, fi, fo):
    torch = torch_cuda
    rng = np.random.default_rng(n + batch)
    x = rng.standard_normal((batch, n))
    z = mc.real_to_pairs(x)
    zq = mc.decode(mc.encode(z, fi), fi)
    xq = mc.pairs_to_real(zq, n)
    p = mc.RealPlan(n, batch, "r2c", fi, fo)
    d_in = mc.to_device(z, fi)
    d_out = mc.empty((n // 2 + 1) * batch, fo)
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    y = p.result(d_out)
    p.destroy()
    ref = oracle.fft(xq.reshape(-1).astype(np.complex128), (n,)).reshape(batch, n)[:, : n // 2 + 1]
    err = oracle.rel_l2(y, ref.reshape(-1))
    print(f"r2c n={n} b={batch} {fi}->{fo} rel_l2={err:.3e}")
    assert err <= TOL[fo if fo == "split" else fi]


@pytest.mark.parametrize("n,batch", [(32, 64), (1024, 32), (65536, 4), (131072, 2)])
@pytest.mark.parametrize("fi,fo", FMTS)
def test_c2r(torch_cuda, n, batch, fi, fo):
    torch = torch_cuda
    rng = np.random.default_rng(n + 3)
    h = n // 2
    spec = rng.standard_normal((batch, h + 1)) + 1j * rng.standard_normal((batch, h + 1))
    spec[:, 0] = spec[:, 0].real
    spec[:, h] = spec[:, h].real
    sq = mc.decode(mc.encode(spec.reshape(-1), fi), fi).reshape(batch, h + 1)
    full = np.concatenate([sq, np.conj(sq[:, h - 1:0:-1])], axis=1)
    ref = oracle.fft(full.reshape(-1), (n,), sign=+1).real.reshape(batch, n)
    p = mc.RealPlan(n, batch, "c2r", fi, fo)
    d_in = mc.to_device(spec.reshape(-1), fi)
    d_out = mc.empty(h * batch, fo)
    p.exec(d_in, d_out, mc.INVERSE)
    torch.cuda.synchronize()
    y = mc.pairs_to_real(p.result(d_out), n)
    p.destroy()
    err = oracle.rel_l2(y.reshape(-1).astype(np.complex128), ref.reshape(-1).astype(np.complex128))
    print(f"c2r n={n} b={batch} {fi}->{fo} rel_l2={err:.3e}")
    assert err <= TOL[fo if fo == "split" else fi]


def test_r2c_then_c2r_round_trip(torch_cuda):
    """c2r(r2c(x)) = n x (split both ways)."""
    torch = torch_cuda
    n, batch = 8192, 8
    x = np.random.default_rng(4).standard_normal((batch, n))
    z = mc.real_to_pairs(x)
    xq = mc.pairs_to_real(mc.decode(mc.encode(z, "fp16"), "fp16"), n)
    pf = mc.RealPlan(n, batch, "r2c", "fp16", "split")
    pi = mc.RealPlan(n, batch, "c2r", "split", "split")
    d_in = mc.to_device(z, "fp16")
    d_mid = mc.empty((n // 2 + 1) * batch, "split")
    d_out = mc.empty(n // 2 * batch, "split")
    pf.exec(d_in, d_mid, mc.FORWARD)
    pi.exec(d_mid, d_out, mc.INVERSE, in_exp=pf)   # the forward's exponents carried through
    torch.cuda.synchronize()
    y = mc.pairs_to_real(pi.result(d_out), n) / n
    pf.destroy()
    pi.destroy()
    assert oracle.rel_l2(y.reshape(-1).astype(np.complex128), xq.reshape(-1).astype(np.complex128)) <= 1e-6


def test_real_plan_exponent_queries_agree(torch_cuda):
    """mcfft_get_item_exponents and mcfft_get_output_exponent report the same
    exponent for real plans (C2R carries one more power of two)."""
    import ctypes
    for kind in ("r2c", "c2r"):
        p = mc.RealPlan(4096, 4, kind, "fp16", "fp16", "balanced")
        e = ctypes.c_int()
        assert mc.lib().mcfft_get_item_exponents(p._h, ctypes.byref(e)) == 0
        assert e.value == p.exponent
        q = mc.Plan([2048], 4, "fp16", "fp16", "balanced")
        assert p.exponent == q.exponent + (1 if kind == "c2r" else 0)
        p.destroy()
        q.destroy()


ND = [([1024, 1024], 2), ([64, 64, 128], 2), ([4096, 512], 1), ([32, 4096], 3)]


@pytest.mark.parametrize("dims,batch", ND)
@pytest.mark.parametrize("fi,fo", FMTS)
def test_r2c_nd(torch_cuda, dims, batch, fi, fo):
    """2D/3D R2C against the FP64 oracle's complex FFT of the real input (the
    non-redundant half of the last dimension)."""
    torch = torch_cuda
    n, h = dims[-1], dims[-1] // 2
    rng = np.random.default_rng(sum(dims) + batch)
    x = rng.standard_normal([batch] + dims)
    z = mc.real_to_pairs(x)
    xq = mc.pairs_to_real(mc.decode(mc.encode(z, fi), fi), n).reshape([batch] + dims)
    p = mc.RealPlan(dims, batch, "r2c", fi, fo)
    d_in = mc.to_device(z, fi)
    d_out = mc.empty(p.output_elems_per_transform() * batch, fo)
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    y = p.result(d_out)
    p.destroy()
    ref = oracle.fft(xq.reshape(-1).astype(np.complex128), tuple(dims)).reshape([batch] + dims)[..., : h + 1]
    err = oracle.rel_l2(y, ref.reshape(-1))
    print(f"r2c {dims} b={batch} {fi}->{fo} rel_l2={err:.3e}")
    assert err <= TOL[fo if fo == "split" else fi]


@pytest.mark.parametrize("dims,batch", ND)
@pytest.mark.parametrize("fi,fo", FMTS)
def test_c2r_nd(torch_cuda, dims, batch, fi, fo):
    """2D/3D C2R of a Hermitian spectrum (the half spectrum of a real array)
    against the oracle's inverse complex FFT of the full quantised spectrum."""
    torch = torch_cuda
    n, h = dims[-1], dims[-1] // 2
    axes = tuple(range(1, len(dims) + 1))
    rng = np.random.default_rng(sum(dims) + 7)
    spec = np.fft.rfftn(rng.standard_normal([batch] + dims), axes=axes) / np.sqrt(np.prod(dims))
    sq = mc.decode(mc.encode(spec.reshape(-1), fi), fi).reshape(spec.shape)
    ref = np.fft.irfftn(sq, s=dims, axes=axes) * np.prod(dims)   # the Hermitian extension's inverse DFT
    p = mc.RealPlan(dims, batch, "c2r", fi, fo)
    d_in = mc.to_device(spec.reshape(-1), fi)
    d_out = mc.empty(p.output_elems_per_transform() * batch, fo)
    p.exec(d_in, d_out, mc.INVERSE)
    torch.cuda.synchronize()
    y = mc.pairs_to_real(p.result(d_out), n)
    p.destroy()
    err = oracle.rel_l2(y.reshape(-1).astype(np.complex128), ref.reshape(-1).astype(np.complex128))
    print(f"c2r {dims} b={batch} {fi}->{fo} rel_l2={err:.3e}")
    assert err <= TOL[fo if fo == "split" else fi]


def test_r2c_c2r_nd_round_trip(torch_cuda):
    """c2r(r2c(x)) = prod(dims) x for a 2D real array, split both ways, exponents
    carried from the forward plan."""
    torch = torch_cuda
    dims, batch = [2048, 1024], 2
    n = dims[-1]
    x = np.random.default_rng(12).standard_normal([batch] + dims)
    z = mc.real_to_pairs(x)
    xq = mc.pairs_to_real(mc.decode(mc.encode(z, "fp16"), "fp16"), n)
    pf = mc.RealPlan(dims, batch, "r2c", "fp16", "split")
    pi = mc.RealPlan(dims, batch, "c2r", "split", "split")
    d_in = mc.to_device(z, "fp16")
    d_mid = mc.empty(pf.output_elems_per_transform() * batch, "split")
    d_out = mc.empty(pi.output_elems_per_transform() * batch, "split")
    pf.exec(d_in, d_mid, mc.FORWARD)
    pi.exec(d_mid, d_out, mc.INVERSE, in_exp=pf)
    torch.cuda.synchronize()
    y = mc.pairs_to_real(pi.result(d_out), n) / np.prod(dims)
    pf.destroy()
    pi.destroy()
    assert oracle.rel_l2(y.reshape(-1).astype(np.complex128), xq.reshape(-1).astype(np.complex128)) <= 1e-6
