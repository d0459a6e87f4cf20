"""GPU: the measured configurations at their full sizes, through a property that
holds at any size: forward then inverse gives N x (unnormalised FFTs), checked
in FP64 on the device (mcfft_error_norms) against the quantised input.

  2D 4096 x 4096 x 4, 3D 512^3, 1D 65536 x 4096, and the 20064-item ragged batch
  (1D N = 2^8..2^18 and 2D sides 256..2048, about 1e9 elements)

Bounds: split (FP16 in, hi+lo through both directions) 2e-6; plain FP16 3e-3
(two roundings per stage over both directions; one direction is held to 1.5e-3
elsewhere)."""
import math

import numpy as np
import pytest

import mcfft_amd as mc

pytestmark = pytest.mark.gpu

TOL = {"split": 2e-6, "fp16": 3e-3}


def _fmts(mode):
    # (forward in, forward out = inverse in, inverse out)
    return ("fp16", "split", "split") if mode == "split" else ("fp16", "fp16", "fp16")


def _input(torch, elems, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.randn((elems, 2), dtype=torch.float16, device="cuda", generator=g)
    return x, x.double()


def norms_by_exponent(y, fmt, exps, n, ref):
    """Worst rel-L2 over the runs of consecutive transforms that share an
    exponent (one FP64 GPU reduction per run); exps: per-transform exponents of
    y's transforms of n elements, the true round trip being y * 2^e / n."""
    worst, t = 0.0, 0
    while t < len(exps):
        u = t
        while u < len(exps) and exps[u] == exps[t]:
            u += 1
        rel, _ = mc.error_norms(y[t * n:u * n], fmt, int(exps[t]) - int(math.log2(n)), ref[2 * t * n:2 * u * n])
        worst = max(worst, rel)
        t = u
    return worst


@pytest.mark.parametrize("dims,batch", [([4096, 4096], 4), ([512, 512, 512], 1), ([65536], 4096)])
@pytest.mark.parametrize("mode", ["split", "fp16"])
@pytest.mark.parametrize("scaling", ["auto", "balanced"])
def test_roundtrip_full_size(torch_cuda, dims, batch, mode, scaling):
    """The inverse takes the forward's per-transform exponents as its input
    exponents, so its own exponents give the true round trip."""
    torch = torch_cuda
    n = int(np.prod(dims))
    print(f"full-size round trip dims={dims} b={batch} {mode} {scaling}", flush=True)
    f_in, mid, f_out = _fmts(mode)
    x, ref = _input(torch, n * batch, seed=len(dims) * 7 + batch)
    pf = mc.Plan(dims, batch, f_in, mid, scaling)
    pi = mc.Plan(dims, batch, mid, f_out, scaling)
    y1, y2 = mc.empty(n * batch, mid), mc.empty(n * batch, f_out)
    pf.exec(x, y1, mc.FORWARD)
    pi.exec(y1, y2, mc.INVERSE, in_exp=pf)
    e = pi.exponents()
    rel = norms_by_exponent(y2, f_out, e, n, ref.reshape(-1))
    print(f"  rel_l2 {rel:.3e} exponents {e.min()}..{e.max()}", flush=True)
    assert pf.status_flags() == 0 and pi.status_flags() == 0
    assert rel <= TOL[mode]
    pf.destroy()
    pi.destroy()


@pytest.mark.parametrize("mode", ["split", "fp16"])
def test_ragged_roundtrip_c5_full_size(torch_cuda, mode):
    """The 20064-item ragged batch (items sorted by shape and back to back, so the
    large 1D runs go to batched plans and the 2D items through the queue), forward
    then inverse, checked per shape group (one exponent per group)."""
    torch = torch_cuda
    rng = np.random.default_rng(5)
    shapes = [[int(2 ** k)] for k in rng.integers(8, 19, 20000)]
    sides = [256, 512, 1024, 2048]
    shapes += [[int(rng.choice(sides)), int(rng.choice(sides))] for _ in range(64)]
    shapes.sort(key=lambda s: (len(s), s))
    sizes = [int(np.prod(s)) for s in shapes]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    total = int(sum(sizes))
    print(f"C5 round trip {mode}: {len(shapes)} items, {total} elements", flush=True)
    f_in, mid, f_out = _fmts(mode)
    x, ref = _input(torch, total, seed=5)
    items = [(s, int(o), int(o)) for s, o in zip(shapes, offs)]
    pf = mc.RaggedPlan(items, f_in, mid)
    pi = mc.RaggedPlan(items, mid, f_out)
    y1, y2 = mc.empty(total, mid), mc.empty(total, f_out)
    pf.exec(x, y1, mc.FORWARD)
    pi.exec(y1, y2, mc.INVERSE)
    torch.cuda.synchronize()
    assert pf.status_flags() == 0 and pi.status_flags() == 0
    worst, groups, i = 0.0, 0, 0
    while i < len(shapes):
        j = i
        while j < len(shapes) and shapes[j] == shapes[i]:
            j += 1
        e = {pf.exponents[k] + pi.exponents[k] for k in range(i, j)}
        assert len(e) == 1, (shapes[i], e)   # one path per shape group
        lo, hi = int(offs[i]), int(offs[j - 1]) + sizes[j - 1]
        rel, _ = mc.error_norms(y2[lo:hi], f_out, e.pop() - int(math.log2(sizes[i])), ref[lo:hi])
        worst = max(worst, rel)
        groups += 1
        i = j
    print(f"  {groups} shape groups, worst rel_l2 {worst:.3e}", flush=True)
    assert worst <= TOL[mode]
    pf.destroy()
    pi.destroy()
