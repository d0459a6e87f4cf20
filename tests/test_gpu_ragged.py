"""GPU: ragged batches (one persistent launch over a device-side work queue) and
hipGraph capture of exec calls, against the FP64 oracle."""
import numpy as np
import pytest

import mcfft_amd as mc
import oracle

pytestmark = pytest.mark.gpu
TOL = {"fp16": 1.5e-3, "split": 1e-6}


def _mixed(seed, n_small=200):
    rng = np.random.default_rng(seed)
    shapes = [[int(2 ** k)] for k in rng.integers(4, 19, n_small)]
    shapes += [[256, 512], [64, 64], [16, 16, 32], [1024, 256], [32, 2048]]
    rng.shuffle(shapes)
    return shapes


@pytest.mark.parametrize("fmt", ["fp16", "split"])
@pytest.mark.parametrize("sorted_layout", [False, True])
def test_ragged_matches_oracle(torch_cuda, fmt, sorted_layout):
    torch = torch_cuda
    shapes = _mixed(1)
    if sorted_layout:
        shapes.sort(key=lambda s: (len(s), int(np.prod(s))))
    sizes = [int(np.prod(s)) for s in shapes]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sum(sizes))
    rng = np.random.default_rng(2)
    x = rng.standard_normal(total) + 1j * rng.standard_normal(total)
    in_fmt, out_fmt = ("fp16", "split") if fmt == "split" else ("fp16", "fp16")
    p = mc.RaggedPlan([(s, o, o) for s, o in zip(shapes, offs)], in_fmt, out_fmt)
    d_in = mc.to_device(x, in_fmt)
    d_out = mc.empty(total, out_fmt)
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    assert p.status_flags() == 0
    y = mc.from_device(d_out, out_fmt)
    xq = mc.decode(mc.encode(x, in_fmt), in_fmt)
    worst = 0.0
    for s, o, e in zip(shapes, offs, p.exponents):
        n = int(np.prod(s))
        err = oracle.rel_l2(y[o:o + n] * 2.0 ** e, oracle.fft(xq[o:o + n], tuple(s)))
        worst = max(worst, err)
    print(f"ragged {fmt} sorted={sorted_layout}: {len(shapes)} items, worst rel_l2 {worst:.3e}")
    assert worst <= TOL[fmt]


@pytest.mark.parametrize("fmt", ["fp16", "split"])
def test_ragged_single_launch_matches_phased(torch_cuda, monkeypatch, fmt):
    """MCFFT_RAGGED_SINGLE=1: every pass in one launch, later-pass tiles waiting on
    their transform's tile counter -- the same tiles and arithmetic as the phased
    launches, so the same bits; no wait may run out (status 0)."""
    torch = torch_cuda
    shapes = _mixed(1)
    sizes = [int(np.prod(s)) for s in shapes]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sum(sizes))
    print(f"ragged single launch {fmt}: {len(shapes)} items", flush=True)
    x = np.random.default_rng(4).standard_normal(total) * (1 + 1j)
    in_fmt, out_fmt = ("fp16", "split") if fmt == "split" else ("fp16", "fp16")
    p = mc.RaggedPlan([(s, o, o) for s, o in zip(shapes, offs)], in_fmt, out_fmt)
    d_in = mc.to_device(x, in_fmt)
    phased, single = mc.empty(total, out_fmt), mc.empty(total, out_fmt)
    p.exec(d_in, phased, mc.FORWARD)
    monkeypatch.setenv("MCFFT_RAGGED_SINGLE", "1")
    p.exec(d_in, single, mc.FORWARD)
    torch.cuda.synchronize()
    assert p.status_flags() == 0
    assert torch.equal(phased.view(torch.int16), single.view(torch.int16))


@pytest.mark.parametrize("bulk", ["default", "0"])
def test_ragged_bulk_runs(torch_cuda, monkeypatch, bulk):
    """Runs of same-shape items lying back to back with >= 2^23 elements run as a
    batched plan each (two here: 70 x 131072 and 40 x 512x512), the other items in
    the persistent queue; MCFFT_RAGGED_BULK=0 puts every item in the queue.  Every
    item against the oracle with its own exponent."""
    torch = torch_cuda
    if bulk != "default":
        monkeypatch.setenv("MCFFT_RAGGED_BULK", bulk)
    shapes = [[131072]] * 70 + [[4096]] * 3 + [[256, 512]] + [[65536]] * 5 + [[512, 512]] * 40 + [[1024]] * 9
    sizes = [int(np.prod(s)) for s in shapes]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sum(sizes))
    print(f"ragged bulk={bulk}: {len(shapes)} items, {total} elements", flush=True)
    x = np.random.default_rng(8).standard_normal(total) + 1j * np.random.default_rng(9).standard_normal(total)
    p = mc.RaggedPlan([(s, o, o) for s, o in zip(shapes, offs)], "fp16", "fp16")
    desc = p.describe()
    assert desc.count("bulk in=") == (2 if bulk == "default" else 0), desc
    d_in = mc.to_device(x, "fp16")
    d_out = mc.empty(total, "fp16")
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    assert p.status_flags() == 0
    y = mc.from_device(d_out, "fp16")
    xq = mc.decode(mc.encode(x, "fp16"), "fp16")
    worst = 0.0
    for i in list(range(0, 70, 23)) + list(range(70, len(shapes))):
        s, o, e = shapes[i], int(offs[i]), p.exponents[i]
        n = int(np.prod(s))
        worst = max(worst, oracle.rel_l2(y[o:o + n] * 2.0 ** e, oracle.fft(xq[o:o + n], tuple(s))))
    print(f"  worst rel_l2 {worst:.3e}", flush=True)
    assert worst <= TOL["fp16"]
    # the exec (batched plans + queue launches) replays from a hipGraph to the same bits
    o2 = mc.empty(total, "fp16")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            p.exec(d_in, o2, mc.FORWARD, stream=s)
    torch.cuda.synchronize()
    for _ in range(3):
        o2.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(o2.view(torch.int16), d_out.view(torch.int16))
    p.destroy()


def test_ragged_inverse_and_gaps(torch_cuda):
    """Items at arbitrary (gapped, out-of-order) offsets, inverse direction."""
    torch = torch_cuda
    shapes = [[4096], [131072], [512], [128, 128]]
    sizes = [int(np.prod(s)) for s in shapes]
    in_offs = [300000, 0, 140000, 150000]
    out_offs = [0, 10000, 150000, 200000]
    rng = np.random.default_rng(5)
    total_in, total_out = 320000, 220000
    x = rng.standard_normal(total_in) + 1j * rng.standard_normal(total_in)
    p = mc.RaggedPlan([(s, i, o) for s, i, o in zip(shapes, in_offs, out_offs)], "fp16", "split")
    d_in = mc.to_device(x, "fp16")
    d_out = torch.zeros((total_out, 4), dtype=torch.float16, device="cuda")
    p.exec(d_in, d_out, mc.INVERSE)
    torch.cuda.synchronize()
    y = mc.from_device(d_out, "split")
    xq = mc.decode(mc.encode(x, "fp16"), "fp16")
    for s, n, i, o, e in zip(shapes, sizes, in_offs, out_offs, p.exponents):
        err = oracle.rel_l2(y[o:o + n] * 2.0 ** e, oracle.fft(xq[i:i + n], tuple(s), sign=+1))
        assert err <= 1e-6, (s, err)


def test_graph_capture_replays(torch_cuda):
    """exec calls captured into a hipGraph (through torch.cuda.graph) give the
    eager result on every replay, with eager work between replays."""
    torch = torch_cuda
    n, b = 131072, 8
    rng = np.random.default_rng(9)
    x = rng.standard_normal(n * b) + 1j * rng.standard_normal(n * b)
    p = mc.Plan([n], b, "fp16", "split")
    pr = mc.RaggedPlan([([4096], 0, 0), ([65536], 4096, 4096), ([256, 256], 69632, 69632)], "fp16", "fp16")
    d_in = mc.to_device(x, "fp16")
    d_out = mc.empty(n * b, "split").zero_()
    d_rout = mc.empty(n * b, "fp16").zero_()   # the ragged items cover only a prefix
    p.exec(d_in, d_out, mc.FORWARD)
    pr.exec(d_in, d_rout, mc.FORWARD)
    torch.cuda.synchronize()
    ref, rref = d_out.clone(), d_rout.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            p.exec(d_in, d_out, mc.FORWARD)
            pr.exec(d_in, d_rout, mc.FORWARD)
    torch.cuda.synchronize()
    for rep in range(4):
        d_out.zero_()
        d_rout.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(d_out, ref), rep
        assert torch.equal(d_rout, rref), rep
    assert pr.status_flags() == 0


@pytest.mark.parametrize("fmt", ["fp16", "split"])
def test_ragged_gathered_bulk_runs(torch_cuda, monkeypatch, fmt):
    """Same-shape items scattered through the batch (drawn order, with gaps)
    run as gathered batched plans when together they reach the bulk size
    (MCFFT_RAGGED_BULK lowered to 2^18 here: the 8192- and 131072-point items);
    items that reach it alone (the 2D and 3D ones) run as one-item batched plans;
    short 1D rows stay in the queue.  Every item against the oracle; nothing
    between items written."""
    torch = torch_cuda
    monkeypatch.setenv("MCFFT_RAGGED_BULK", str(1 << 18))
    rng = np.random.default_rng(9)
    shapes = [[8192]] * 40 + [[1 << 17]] * 3 + [[1024, 512]] * 2 + [[64, 64, 64]] * 2 + [[256]] * 30
    rng.shuffle(shapes)
    sizes = [int(np.prod(s)) for s in shapes]
    gaps = rng.integers(0, 64, len(shapes))
    offs, o = [], 0
    for n, g in zip(sizes, gaps):
        o += int(g)
        offs.append(o)
        o += n
    total = o + 16
    x = rng.standard_normal(total) + 1j * rng.standard_normal(total)
    in_fmt, out_fmt = ("fp16", "split") if fmt == "split" else ("fp16", "fp16")
    p = mc.RaggedPlan([(s, q, q) for s, q in zip(shapes, offs)], in_fmt, out_fmt)
    desc = p.describe()
    assert desc.count("bulk gathered") == 2 and desc.count("  bulk") == 6, desc
    d_in = mc.to_device(x, in_fmt)
    d_out = mc.to_device(np.full(total, 7.0 + 0j), out_fmt)
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    assert p.status_flags() == 0
    y = mc.from_device(d_out, out_fmt)
    xq = mc.decode(mc.encode(x, in_fmt), in_fmt)
    worst, covered = 0.0, np.zeros(total, dtype=bool)
    for s, q, e in zip(shapes, offs, p.exponents):
        n = int(np.prod(s))
        covered[q:q + n] = True
        worst = max(worst, oracle.rel_l2(y[q:q + n] * 2.0 ** e, oracle.fft(xq[q:q + n], tuple(s))))
    print(f"ragged gathered {fmt}: {len(shapes)} items, worst rel_l2 {worst:.3e}")
    assert worst <= TOL[fmt]
    assert np.all(y[~covered] == 7.0)
    # the exec (gathered and one-item batched plans + queue launches) replays from
    # a hipGraph to the same bits, eager work and synchronisations in between
    o2 = mc.to_device(np.full(total, 7.0 + 0j), out_fmt)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            p.exec(d_in, o2, mc.FORWARD, stream=s)
    torch.cuda.synchronize()
    for _ in range(3):
        o2.zero_()
        g.replay()
        torch.cuda.synchronize()
        got = mc.from_device(o2, out_fmt)
        assert np.array_equal(got[covered], y[covered])


def test_ragged_many_short_items_gathered(torch_cuda):
    """1024 or more items of one 1D length whose row pass has a one-row tile
    (2048 and 4096 points here, shuffled among 1024-point items, with gaps) run
    as gathered batched plans although they hold fewer than 2^23 elements;
    1024-point items, whose tiles hold several rows, stay in the queue.  A
    sample of every shape against the oracle; nothing between items written."""
    torch = torch_cuda
    rng = np.random.default_rng(31)
    shapes = [[2048]] * 1100 + [[4096]] * 1030 + [[1024]] * 1200
    rng.shuffle(shapes)
    sizes = [int(np.prod(s)) for s in shapes]
    gaps = rng.integers(0, 16, len(shapes))
    offs, o = [], 0
    for n, g in zip(sizes, gaps):
        o += int(g)
        offs.append(o)
        o += n
    total = o + 16
    x = rng.standard_normal(total) + 1j * rng.standard_normal(total)
    p = mc.RaggedPlan([(s, q, q) for s, q in zip(shapes, offs)], "fp16", "fp16")
    desc = p.describe()
    assert desc.count("bulk gathered") == 2, desc
    assert "dims=2048 batch=1100" in desc and "dims=4096 batch=1030" in desc, desc
    d_in = mc.to_device(x, "fp16")
    d_out = mc.to_device(np.full(total, 7.0 + 0j), "fp16")
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    assert p.status_flags() == 0
    y = mc.from_device(d_out, "fp16")
    xq = mc.decode(mc.encode(x, "fp16"), "fp16")
    covered = np.zeros(total, dtype=bool)
    for s, q in zip(shapes, offs):
        covered[q:q + int(np.prod(s))] = True
    assert np.all(y[~covered] == 7.0)
    worst = 0.0
    for n in (1024, 2048, 4096):
        idx = [i for i, s in enumerate(shapes) if s[0] == n]
        for i in rng.choice(idx, 12, replace=False):
            q, e = offs[i], p.exponents[i]
            worst = max(worst, oracle.rel_l2(y[q:q + n] * 2.0 ** e, oracle.fft(xq[q:q + n], (n,))))
    print(f"ragged short items: {len(shapes)} items, worst sampled rel_l2 {worst:.3e}")
    assert worst <= TOL["fp16"]


def test_ragged_shapes_outside_the_queue(torch_cuda):
    """Shapes the queue's pass variants cannot run -- narrow outer axes (256 x 16,
    512 x 16, 128 x 64 x 16) and a 1D length above 2^18 -- run as gathered
    batched plans inside the same ragged exec, beside queue items."""
    torch = torch_cuda
    shapes = [[256, 16], [1024], [512, 16], [1 << 20], [256, 16], [128, 64, 16], [4096], [64]]
    sizes = [int(np.prod(s)) for s in shapes]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sum(sizes))
    rng = np.random.default_rng(21)
    x = rng.standard_normal(total) + 1j * rng.standard_normal(total)
    p = mc.RaggedPlan([(s, o, o) for s, o in zip(shapes, offs)], "fp16", "fp16")
    desc = p.describe()
    assert desc.count("bulk gathered") == 4, desc
    d_in = mc.to_device(x, "fp16")
    d_out = mc.empty(total, "fp16")
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    assert p.status_flags() == 0
    y = mc.from_device(d_out, "fp16")
    xq = mc.decode(mc.encode(x, "fp16"), "fp16")
    worst = max(oracle.rel_l2(y[o:o + n] * 2.0 ** e, oracle.fft(xq[o:o + n], tuple(s)))
                for s, o, n, e in zip(shapes, offs, sizes, p.exponents))
    assert worst <= TOL["fp16"], worst
