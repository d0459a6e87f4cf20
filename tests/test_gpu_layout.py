"""GPU: strided / planar layouts (mcfft_plan_many_layout, mcfft_exec_layout).

A layout exec must equal the contiguous exec of the same values bit for bit
(the gather and scatter kernels only move bits and the transform kernels are
the same), match the FP64 oracle, and leave output positions outside the
layout untouched.  The full-size case (1D 131072 x 256, planar on both sides)
checks the bit-for-bit property only."""
import numpy as np
import pytest

import emu  # host-side layout encoding helpers (pure numpy)
import mcfft_amd as mc
import oracle

pytestmark = pytest.mark.gpu
TOL = {"fp16": 1.5e-3, "bf16": 1.2e-2, "split": 1e-6}
FMT_ID = {"fp16": emu.FMT_F16, "bf16": emu.FMT_BF16, "split": emu.FMT_SPLIT}
FILL = 0x7C01


def _dense(dims):
    s, st = 1, []
    for d in reversed(dims):
        st.append(s)
        s *= d
    return list(reversed(st)), s


CASES = [
    ([256], 3, ([3], 3 * 256 + 5, 0), None, "fp16", "fp16"),
    ([512], 2, None, ([2], 1100, 1), "bf16", "bf16"),
    ([64, 32], 2, ([1, 64], 64 * 32, 1), ([1, 64], 64 * 32, 1), "fp16", "fp16"),
    ([16, 16, 32], 1, ([16 * 40, 40, 1], 16 * 16 * 40, 0), ([16 * 40, 40, 1], 16 * 16 * 40, 1), "fp16", "fp16"),
    ([1024], 3, ([1], 1024, 1), ([5], 5 * 1024 + 1, 1), "fp16", "split"),
    ([2048], 2, ([2], 4097, 0), ([1], 2048, 1), "split", "split"),
    ([131072], 3, ([1], 131072, 1), ([1], 131072, 1), "fp16", "fp16"),          # four-step, planar
    ([4096, 64], 2, ([1, 4096], 4096 * 64, 0), None, "bf16", "bf16"),              # column-major input, 2D
    ([65536], 2, ([1], 65536, 1), ([1], 65536, 1), "split", "split"),            # split four-step, shared staging
    # spaced transforms (dense inside, a gap between): no staging, the passes
    # locate transform t at t * dist
    ([1024, 64], 2, ([64, 1], 1024 * 64 + 96, 0), ([64, 1], 1024 * 64 + 32, 0), "fp16", "fp16"),
    ([131072], 3, ([1], 131072 + 64, 0), ([1], 131072 + 8, 0), "fp16", "split"),   # four-step
    ([16384], 4, ([1], 16384 + 7, 0), None, "bf16", "bf16"),                        # one-row tiles
    ([1024], 3, ([1], 1030, 0), None, "fp16", "fp16"),                              # 4-row tiles: staged
    ([4096], 5, ([1], 4096 + 64, 0), ([1], 4096 + 3, 0), "fp16", "fp16"),          # one-row tile kernel when gathered
    ([8192], 3, None, ([1], 8192 + 16, 0), "bf16", "bf16"),
]
SPACED = {9: True, 10: True, 11: True, 12: False, 13: True, 14: True}


def _to_dev(torch, a):
    return None if a is None else torch.from_numpy(a.view(np.int16)).cuda()


def _from_dev(t):
    return None if t is None else t.cpu().numpy().view(np.uint16)


def run_layout(torch, x, dims, batch, li, lo, fi, fo, direction):
    st, n = _dense(dims)
    li = li or (st, n, 0)
    lo = lo or (st, n, 0)
    ipos = emu.layout_positions(dims, batch, li[0], li[1])
    opos = emu.layout_positions(dims, batch, lo[0], lo[1])
    in_re, in_im = emu.layout_encode(x, FMT_ID[fi], li[2], ipos, int(ipos.max()) + 1)
    out_re, out_im = emu.layout_encode(np.zeros(0), FMT_ID[fo], lo[2], np.zeros(0, dtype=np.int64),
                                       int(opos.max()) + 1, fill=FILL)
    p = mc.LayoutPlan(dims, batch, mc.Layout.make(*li), mc.Layout.make(*lo), fi, fo)
    run_layout.staging_bytes = p.workspace_bytes() - mc.Plan(dims, batch, fi, fo).workspace_bytes()
    d = [_to_dev(torch, a) for a in (in_re, in_im, out_re, out_im)]
    p.exec_layout(d[0], d[1], d[2], d[3], direction)
    torch.cuda.synchronize()
    ore, oim = _from_dev(d[2]), _from_dev(d[3])
    e = p.exponents()   # per transform (AUTO by default)
    y = (emu.layout_decode(ore, oim, FMT_ID[fo], lo[2], opos).reshape(batch, -1) *
         np.exp2(e.astype(np.float64))[:, None]).reshape(-1)
    p.destroy()
    return y, (ore, oim), opos


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("direction", [mc.FORWARD, mc.INVERSE])
def test_layout_exec(torch_cuda, case, direction):
    torch = torch_cuda
    dims, batch, li, lo, fi, fo = CASES[case]
    n = int(np.prod(dims))
    print(f"layout case {case} dims={dims} b={batch} {fi}->{fo} dir={direction}", flush=True)
    rng = np.random.default_rng(case)
    x = rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)
    y, (ore, oim), opos = run_layout(torch, x, dims, batch, li, lo, fi, fo, direction)
    if case in SPACED:
        assert (run_layout.staging_bytes == 0) == SPACED[case], run_layout.staging_bytes
    # contiguous exec of the same values
    p = mc.Plan(dims, batch, fi, fo)
    d_in = mc.to_device(x, fi)
    d_out = mc.empty(n * batch, fo)
    p.exec(d_in, d_out, direction)
    torch.cuda.synchronize()
    y0 = p.result(d_out)
    p.destroy()
    assert np.array_equal(y, y0), case
    xq = mc.decode(mc.encode(x, fi), fi)
    ref = oracle.fft(xq, tuple(dims), sign=+1 if direction == mc.INVERSE else -1)
    err = oracle.rel_l2(y, ref)
    print(f"  rel_l2={err:.3e}")
    assert err <= TOL[fo if fo == "split" else fi]
    for a in (ore, oim):
        if a is not None:
            w = a.reshape(int(opos.max()) + 1, -1)
            mask = np.ones(w.shape[0], dtype=bool)
            mask[opos] = False
            assert np.all(w[mask] == FILL)


def test_layout_full_size_planar(torch_cuda):
    """1D 131072 x 256, FP16, planar input and output: bit for bit the
    contiguous exec (size-independent check)."""
    torch = torch_cuda
    n, batch = 131072, 256
    g = torch.Generator(device="cuda").manual_seed(5)
    inter = torch.randn((n * batch, 2), device="cuda", generator=g).half()
    re, im = inter[:, 0].contiguous(), inter[:, 1].contiguous()
    lay = mc.Layout.make([1], n, True)
    p = mc.LayoutPlan([n], batch, lay, lay, "fp16", "fp16")
    o_re, o_im = torch.empty_like(re), torch.empty_like(im)
    p.exec_layout(re, im, o_re, o_im, mc.FORWARD)
    q = mc.Plan([n], batch, "fp16", "fp16")
    out = torch.empty_like(inter)
    q.exec(inter, out, mc.FORWARD)
    torch.cuda.synchronize()
    assert torch.equal(o_re.view(torch.int16), out[:, 0].contiguous().view(torch.int16))
    assert torch.equal(o_im.view(torch.int16), out[:, 1].contiguous().view(torch.int16))
    p.destroy()
    q.destroy()


def test_layout_invalid(torch_cuda):
    torch = torch_cuda
    with pytest.raises(mc.McfftError) as e:
        mc.LayoutPlan([256], 1, mc.Layout.make([0], 256), None)
    assert e.value.status == 1
    p = mc.LayoutPlan([256], 1, mc.Layout.make([1], 256, True), None)
    buf = torch.zeros((256, 2), dtype=torch.float16, device="cuda")
    with pytest.raises(mc.McfftError) as e:   # planar input without its imaginary array
        p.exec_layout(buf, None, buf, None)
    assert e.value.status == 1
    p.destroy()


def test_spaced_layout_in_place_and_overlap(torch_cuda):
    """A spaced layout (no staging) in place gives the out-of-place result bit
    for bit; sides that overlap other than exactly are refused."""
    torch = torch_cuda
    dims, batch, dist = [2048, 32], 3, 2048 * 32 + 64
    n = 2048 * 32
    lay = mc.Layout.make([32, 1], dist)
    p = mc.LayoutPlan(dims, batch, lay, lay, "fp16", "fp16")
    g = torch.Generator(device="cuda").manual_seed(8)
    ext = (batch - 1) * dist + n
    big = torch.randn((ext + 64, 2), device="cuda", generator=g).half()
    buf = big[:ext]
    out = torch.empty_like(buf)
    p.exec_layout(buf, None, out, None, mc.FORWARD)
    p.exec_layout(buf, None, buf, None, mc.FORWARD)   # in place
    torch.cuda.synchronize()
    for t in range(batch):
        sl = slice(t * dist, t * dist + n)
        assert torch.equal(buf[sl].view(torch.int32), out[sl].view(torch.int32))
    with pytest.raises(mc.McfftError) as e:
        p.exec_layout(buf, None, big[64:], None, mc.FORWARD)   # overlaps, not exactly
    assert e.value.status == 1
    p.destroy()
