"""GPU: mcfft_exec_timed (each pass's kernel bracketed by its own events) gives
the same bits as mcfft_exec, and per-pass durations that fit inside the exec."""
import numpy as np
import pytest

import mcfft_amd as mc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,batch,fi,fo", [([131072], 16, "fp16", "fp16"), ([4096], 64, "fp16", "split"),
                                              ([256, 512], 2, "bf16", "bf16")])
def test_exec_timed_matches_exec(torch_cuda, dims, batch, fi, fo):
    torch = torch_cuda
    n = int(np.prod(dims))
    print(f"exec_timed dims={dims} b={batch} {fi}->{fo}", flush=True)
    x = np.random.default_rng(3).standard_normal(n * batch) * (1 + 1j)
    p = mc.Plan(dims, batch, fi, fo)
    d_in = mc.to_device(x, fi)
    o1, o2 = mc.empty(n * batch, fo), mc.empty(n * batch, fo)
    npass = p.pass_count()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(npass)]
    stops = [torch.cuda.Event(enable_timing=True) for _ in range(npass)]
    for e in starts + stops:
        e.record()
    outer0, outer1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    p.exec(d_in, o1, mc.FORWARD)
    outer0.record()
    p.exec_timed(d_in, o2, mc.FORWARD, starts, stops)
    outer1.record()
    torch.cuda.synchronize()
    assert torch.equal(o1.view(torch.int16), o2.view(torch.int16))
    per = [a.elapsed_time(b) for a, b in zip(starts, stops)]
    total = outer0.elapsed_time(outer1)
    print("  per-pass ms", [round(v, 4) for v in per], "exec ms", round(total, 4))
    assert all(v > 0 for v in per)
    assert sum(per) <= total * 1.05 + 0.002
    p.destroy()


def test_exec_timed_rejects_chunked_plans(torch_cuda, monkeypatch):
    torch = torch_cuda
    monkeypatch.setenv("MCFFT_WS_BYTES", str(1 << 20))   # forces several batch chunks
    p = mc.Plan([131072], 8, "fp16", "fp16")
    buf = mc.empty(131072 * 8, "fp16")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(p.pass_count())]
    for e in ev:
        e.record()
    with pytest.raises(mc.McfftError) as e:
        p.exec_timed(buf, buf, mc.FORWARD, ev, ev)
    assert e.value.status == 2
    p.destroy()


def test_python_binding_rejects_undersized_tensors(torch_cuda):
    """A tensor smaller than the plan's input or output is refused before any
    kernel is launched (the C ABI itself only sees pointers)."""
    torch = torch_cuda
    p = mc.Plan([4096], 8, "fp16", "split")
    good_in, good_out = mc.empty(4096 * 8, "fp16"), mc.empty(4096 * 8, "split")
    with pytest.raises(ValueError, match="output"):
        p.exec(good_in, mc.empty(4096 * 8, "fp16"))          # split output needs 8 B per element
    with pytest.raises(ValueError, match="input"):
        p.exec(good_in[:100], good_out)
    with pytest.raises(ValueError, match="contiguous"):
        p.exec(good_in.t(), good_out)
    p.exec(good_in, good_out)
    torch.cuda.synchronize()
    r = mc.RealPlan(4096, 4, "r2c", "fp16", "fp16")
    with pytest.raises(ValueError, match="output"):
        r.exec(mc.empty(2048 * 4, "fp16"), mc.empty(2048 * 4, "fp16"))   # needs 2049 per transform
    p.destroy()
    r.destroy()
