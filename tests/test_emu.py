"""The library's kernels and planner, compiled for the CPU (tests/emu), against
the FP64 oracle -- CPU only.  Every LDS and global access of the emulated
kernels is bounds-checked (an out-of-range index aborts the test process), and
the MFMA is evaluated from the operand layout the GPU layout test pins."""
import ctypes

import numpy as np
import pytest

import emu
import oracle

TOL = {emu.FMT_F16: 1.5e-3, emu.FMT_BF16: 1.2e-2, emu.FMT_SPLIT: 1e-6}
FMTS = [(emu.FMT_F16, emu.FMT_F16), (emu.FMT_BF16, emu.FMT_BF16), (emu.FMT_F16, emu.FMT_SPLIT)]


def check(dims, batch, fi, fo, inverse=False, scaling=1, ws_limit=0, seed=0):
    n = int(np.prod(dims))
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)
    xq = emu.from_storage(emu.to_storage(x, fi), fi)
    y, e = emu.exec_fft(x, list(dims), batch, fi, fo, scaling=scaling, inverse=inverse, ws_limit=ws_limit)
    ref = oracle.fft(xq, tuple(dims), sign=+1 if inverse else -1)
    err = oracle.rel_l2(y, ref)
    assert err <= TOL[fo if fo == emu.FMT_SPLIT else fi], (dims, batch, fi, fo, err)
    return err


@pytest.mark.parametrize("n,batch", [(16, 128), (32, 64), (64, 32), (128, 16), (256, 8), (512, 4), (1024, 2),
                                     (2048, 1), (4096, 1), (256, 5), (16, 3), (8192, 1), (16384, 1)])
@pytest.mark.parametrize("fi,fo", FMTS)
def test_single_pass(n, batch, fi, fo):
    check([n], batch, fi, fo)


def test_single_pass_32768():
    check([32768], 1, emu.FMT_F16, emu.FMT_F16)
    check([32768], 1, emu.FMT_BF16, emu.FMT_BF16)


@pytest.mark.parametrize("fi", [emu.FMT_F16, emu.FMT_BF16])
@pytest.mark.parametrize("inverse", [False, True])
def test_single_pass_65536(fi, inverse):
    """65536 points in one workgroup (registers + two-round LDS exchange); 5
    transforms on 3 workgroups, so two workgroups prefetch a second transform."""
    assert "mcfft_pass_big" in emu.describe([65536], 5, fi, fi)
    check([65536], 5, fi, fi, inverse=inverse)


def test_single_pass_65536_8_waves():
    """The 8-wave variant (128 elements per lane, round taken from bit 1 of the tile)."""
    emu.set_kernel_override("mcfft_pass_big_0")
    try:
        check([65536], 5, emu.FMT_F16, emu.FMT_F16)   # 3 workgroups: two prefetch a second transform
        assert emu.last_kernels().split() == ["mcfft_pass_big_0"]
        check([65536], 3, emu.FMT_F16, emu.FMT_F16, inverse=True)
    finally:
        emu.set_kernel_override(None)
    emu.set_kernel_override("mcfft_pass_big_1")
    try:
        check([65536], 2, emu.FMT_BF16, emu.FMT_BF16)
    finally:
        emu.set_kernel_override(None)


def test_rows_65536_in_2d():
    assert "mcfft_pass_big16_1" in emu.describe([16, 65536], 1, emu.FMT_BF16, emu.FMT_BF16)
    check([16, 65536], 1, emu.FMT_BF16, emu.FMT_BF16)


@pytest.mark.parametrize("n,batch,ws", [(65536, 1, 0), (65536, 3, 65536 * 8), (1 << 17, 1, 0)])
@pytest.mark.parametrize("fi,fo", [(emu.FMT_F16, emu.FMT_F16), (emu.FMT_F16, emu.FMT_SPLIT)])
def test_four_step(n, batch, ws, fi, fo):
    check([n], batch, fi, fo, ws_limit=ws)


@pytest.mark.parametrize("n,batch,fi,fo,scaling,inverse", [
    (1 << 15, 2, emu.FMT_F16, emu.FMT_SPLIT, 1, False),
    (1 << 15, 1, emu.FMT_F16, emu.FMT_SPLIT, 3, True),
    (1 << 17, 2, emu.FMT_F16, emu.FMT_F16, 3, False),
    (1 << 17, 1, emu.FMT_F16, emu.FMT_F16, 2, True),
    (1 << 18, 1, emu.FMT_BF16, emu.FMT_BF16, 3, False)])
def test_three_pass(monkeypatch, n, batch, fi, fo, scaling, inverse):
    """The three-pass 1D plan (COLF -> COLM -> ROWT, the plan of 2^24..2^27),
    forced down to 2^15..2^17 so that it runs here."""
    monkeypatch.setenv("MCFFT_THREE_PASS_MIN_LG", "15")
    d = emu.describe([n], batch, fi, fo, scaling)
    assert d.count("pass ") == 3 and "_4 " in d.splitlines()[2], d
    check([n], batch, fi, fo, inverse=inverse, scaling=scaling, seed=n + batch)


@pytest.mark.parametrize("fi,fo", [(emu.FMT_F16, emu.FMT_F16), (emu.FMT_F16, emu.FMT_SPLIT)])
def test_three_pass_reconcile(monkeypatch, fi, fo):
    """AUTO, three passes: first-pass tiles (32 consecutive columns m of the
    N1 x M view) with input maxima 2^4..2^8 apart; the middle pass maps each of
    its elements (column n3, row n2) back to first-pass column n2*N3 + n3."""
    monkeypatch.setenv("MCFFT_THREE_PASS_MIN_LG", "15")
    n, batch = 1 << 17, 2
    lg = 17
    lg3 = lg // 3
    lg1 = (lg - lg3 + 1) // 2
    m_cols = n >> lg1
    rng = np.random.default_rng(23)
    scale = np.exp2(4.0 * (((np.arange(n) % m_cols) // 32) % 3))
    x = (rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)) * np.tile(scale, batch)
    y, e, ref = run_auto(x, [n], batch, fi, fo)
    assert oracle.rel_l2(y, ref) <= TOL[fo if fo == emu.FMT_SPLIT else fi]


def test_three_pass_limits():
    """Plans: 2^24..2^27 three passes (split four-step table in two factors),
    2^23 two passes, 2^28 refused."""
    for lg in (24, 25, 26, 27):
        for fi, fo in FMTS:
            d = emu.describe([1 << lg], 1, fi, fo, 3)
            assert d.count("pass ") == 3, (lg, d)
    assert emu.describe([1 << 23], 1, emu.FMT_F16, emu.FMT_F16, 3).count("pass ") == 2
    with pytest.raises(RuntimeError):
        emu.describe([1 << 28], 1, emu.FMT_F16, emu.FMT_F16, 3)


@pytest.mark.parametrize("dims,batch", [((64, 32), 2), ((16, 256), 1), ((256, 64), 1), ((16, 16, 32), 1),
                                        ((32, 64, 32), 1), ((512, 64), 1), ((512, 16, 32), 1)])
def test_multi_dim(dims, batch):
    check(dims, batch, emu.FMT_F16, emu.FMT_F16)
    check(dims, batch, emu.FMT_F16, emu.FMT_SPLIT)


@pytest.mark.parametrize("dims,batch,fo", [((256, 16), 4, emu.FMT_F16), ((256, 32), 2, emu.FMT_F16),
                                           ((256, 32), 2, emu.FMT_SPLIT), ((256, 16), 4, emu.FMT_SPLIT), ((512, 16), 2, emu.FMT_F16),
                                           ((32, 16), 8, emu.FMT_F16), ((64, 16), 4, emu.FMT_SPLIT), ((128, 16), 2, emu.FMT_F16),
                                           ((128, 64, 16), 1, emu.FMT_SPLIT),
                                           ((256, 256, 16), 1, emu.FMT_F16)])
def test_narrow_outer_axes(dims, batch, fo):
    """Outer axes with fewer columns than the preferred column tile run on a
    compiled variant with a narrower tile (16 or 32 columns)."""
    check(dims, batch, emu.FMT_F16, fo)


@pytest.mark.parametrize("dims", [(256,), (512,), (8192,), (64, 64)])
def test_inverse(dims):
    check(dims, 1, emu.FMT_F16, emu.FMT_SPLIT, inverse=True)


@pytest.mark.parametrize("scaling", [0, 1, 2])
def test_scaling_modes(scaling):
    check([1024], 2, emu.FMT_F16, emu.FMT_SPLIT, scaling=scaling)
    check([8192], 1, emu.FMT_F16, emu.FMT_F16, scaling=scaling)


def test_split_input_roundtrip():
    n = 2048
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    xq = emu.from_storage(emu.to_storage(x, emu.FMT_F16), emu.FMT_F16)
    y, e1 = emu.exec_fft(xq, [n], 1, emu.FMT_F16, emu.FMT_SPLIT)
    z, e2 = emu.exec_fft(y, [n], 1, emu.FMT_SPLIT, emu.FMT_SPLIT, inverse=True)
    assert oracle.rel_l2(z / n, xq) < 1e-6


def test_plan_describe_and_errors():
    d = emu.describe([131072], 256, emu.FMT_F16, emu.FMT_F16)
    assert "fourstep" in d and d.count("pass ") == 2
    assert emu.describe([32768], 4, emu.FMT_F16, emu.FMT_F16).count("pass ") == 1   # single pass
    assert emu.describe([32768], 4, emu.FMT_F16, emu.FMT_SPLIT).count("pass ") == 2  # split: four-step
    with pytest.raises(RuntimeError, match="status 2"):
        emu.describe([1000], 1, 0, 0)        # not a power of two
    with pytest.raises(RuntimeError, match="status 2"):
        emu.describe([8], 1, 0, 0)           # below 16
    with pytest.raises(RuntimeError, match="status 2"):
        emu.describe([8192, 16], 1, 0, 0)    # strided side above 4096
    with pytest.raises(RuntimeError, match="status 2"):
        emu.describe([256], 1, 2, 0)         # split in, fp16 out
    with pytest.raises(RuntimeError, match="status 2"):
        emu.describe([16, 1 << 24], 1, 0, 0)  # innermost side above 2^23
    # split rows of 65536 (the single-pass kernel is FP16/BF16 only): four-step rows, then columns
    assert emu.describe([16, 65536], 1, 0, 2).count("pass ") == 3
    with pytest.raises(RuntimeError, match="status 1"):
        emu.describe([256], 0, 0, 0)         # batch 0


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("fi,fo", [(emu.FMT_F16, emu.FMT_F16), (emu.FMT_F16, emu.FMT_SPLIT)])
def test_ragged_queue(fi, fo, single):
    """Mixed lengths/ranks through the persistent ragged kernel; workgroups run
    concurrently, so second-pass tasks really wait on first-pass counters."""
    rng = np.random.default_rng(3)
    shapes = [[256], [256], [256], [8192], [64, 32], [4096], [512], [512], [16384], [16, 16, 32], [32], [2048],
              [128], [65536], [16], [16]]
    xs = [rng.standard_normal(int(np.prod(s))) + 1j * rng.standard_normal(int(np.prod(s))) for s in shapes]
    ys, status = emu.ragged_exec(xs, shapes, fi, fo, nblocks=4, single=single)
    assert status == 0
    for x, s, y in zip(xs, shapes, ys):
        xq = emu.from_storage(emu.to_storage(x, fi), fi)
        err = oracle.rel_l2(y, oracle.fft(xq, tuple(s)))
        assert err <= TOL[fo if fo == emu.FMT_SPLIT else fi], (s, err)


def test_ragged_long_columns():
    """2D items whose strided side is 512 or 2048 (column variants of the ragged
    kernel not reached by the mixed set above)."""
    rng = np.random.default_rng(6)
    shapes = [[2048, 64], [512, 128], [64, 2048]]
    xs = [rng.standard_normal(int(np.prod(s))) + 1j * rng.standard_normal(int(np.prod(s))) for s in shapes]
    ys, status = emu.ragged_exec(xs, shapes, emu.FMT_F16, emu.FMT_F16, nblocks=4, single=True)
    assert status == 0
    for x, s, y in zip(xs, shapes, ys):
        xq = emu.from_storage(emu.to_storage(x, emu.FMT_F16), emu.FMT_F16)
        assert oracle.rel_l2(y, oracle.fft(xq, tuple(s))) <= TOL[emu.FMT_F16], s


def test_ragged_plan_structure_large():
    """Structural checks of the plan of a 20064-item batch (20000 1D lengths
    2^8..2^18 plus 64 2D shapes with sides 256..2048, grouped by shape): every
    dependency wait is satisfiable by earlier queue entries, tiles and buffer
    views are in range (no kernel runs)."""
    rng = np.random.default_rng(5)
    shapes = [[int(2 ** k)] for k in rng.integers(8, 19, 20000)]
    shapes += [[int(rng.choice([256, 512, 1024, 2048])), int(rng.choice([256, 512, 1024, 2048]))] for _ in range(64)]
    shapes.sort(key=lambda s: (len(s), s))
    for fi, fo in ((emu.FMT_F16, emu.FMT_F16), (emu.FMT_F16, emu.FMT_SPLIT)):
        bad, stats = emu.ragged_check(shapes, fi, fo)
        assert bad == 0, stats
        assert stats["phases"] == 2 and stats["tasks"] > 200000


def test_ragged_offsets_and_inverse():
    rng = np.random.default_rng(4)
    shapes = [[1024], [8192], [32, 64]]
    layout = [(20000, 0), (0, 30000), (9000, 1024)]
    xs = [rng.standard_normal(int(np.prod(s))) + 1j * rng.standard_normal(int(np.prod(s))) for s in shapes]
    ys, status = emu.ragged_exec(xs, shapes, emu.FMT_F16, emu.FMT_SPLIT, layout=layout, inverse=True, nblocks=3)
    assert status == 0
    for x, s, y in zip(xs, shapes, ys):
        xq = emu.from_storage(emu.to_storage(x, emu.FMT_F16), emu.FMT_F16)
        assert oracle.rel_l2(y, oracle.fft(xq, tuple(s), sign=+1)) <= 1e-6


@pytest.mark.parametrize("fi,fo", [(emu.FMT_F16, emu.FMT_F16), (emu.FMT_F16, emu.FMT_SPLIT)])
def test_multi_dim_chunked(fi, fo):
    """2D plans run transform chunks (sized to stay cache resident); force one
    transform per chunk and a partial last chunk."""
    check([64, 32], 3, fi, fo, ws_limit=64 * 32 * 8)
    check([32, 16, 32], 2, fi, fo, ws_limit=1)


@pytest.mark.parametrize("n,batch", [(32, 4), (256, 3), (8192, 2), (131072, 1)])
@pytest.mark.parametrize("fi,fo", FMTS)
def test_r2c(n, batch, fi, fo):
    """Real-input forward: X[k], k = 0..n/2, of the complex FFT of the real input."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal((batch, n))
    xq = emu.from_storage(emu.to_storage((x[:, 0::2] + 1j * x[:, 1::2]).reshape(-1), fi), fi).reshape(batch, n // 2)
    xr = np.empty((batch, n))
    xr[:, 0::2], xr[:, 1::2] = xq.real, xq.imag
    y = emu.real_exec("r2c", x, n, batch, fi, fo)
    ref = oracle.fft(xr.reshape(-1).astype(np.complex128), (n,)).reshape(batch, n)[:, : n // 2 + 1]
    err = oracle.rel_l2(y.reshape(-1), ref.reshape(-1))
    assert err <= TOL[fo if fo == emu.FMT_SPLIT else fi], (n, err)


@pytest.mark.parametrize("n,batch", [(32, 4), (1024, 2), (65536, 1)])
@pytest.mark.parametrize("fi,fo", FMTS)
def test_c2r(n, batch, fi, fo):
    """Real-output inverse of a Hermitian spectrum given by its first n/2 + 1 terms."""
    rng = np.random.default_rng(n + 1)
    h = n // 2
    spec = rng.standard_normal((batch, h + 1)) + 1j * rng.standard_normal((batch, h + 1))
    spec[:, 0] = spec[:, 0].real
    spec[:, h] = spec[:, h].real
    sq = emu.from_storage(emu.to_storage(spec.reshape(-1), fi), fi).reshape(batch, h + 1)
    full = np.concatenate([sq, np.conj(sq[:, h - 1:0:-1])], axis=1)
    ref = oracle.fft(full.reshape(-1), (n,), sign=+1).reshape(batch, n)
    assert np.abs(ref.imag).max() <= 1e-9 * np.abs(ref).max()
    y = emu.real_exec("c2r", spec, n, batch, fi, fo)
    err = oracle.rel_l2(y.reshape(-1).astype(np.complex128), ref.real.reshape(-1).astype(np.complex128))
    assert err <= TOL[fo if fo == emu.FMT_SPLIT else fi], (n, err)


@pytest.mark.parametrize("dims,batch", [([16, 64], 2), ([64, 32], 1), ([16, 16, 64], 2), ([16, 32, 32], 1),
                                        ([512, 256], 1)])
@pytest.mark.parametrize("fi,fo", FMTS)
def test_real_nd(dims, batch, fi, fo):
    """2D/3D R2C against the oracle's complex FFT of the real input (the first
    n/2 + 1 bins of the last dimension), and C2R of that quantised half spectrum
    against numpy's irfftn (the inverse DFT of its Hermitian extension): the
    mirror-row pairing (-m, n/2 - k) over every outer dimension, odd and even
    rows, DC and Nyquist bins."""
    n, h = dims[-1], dims[-1] // 2
    axes = tuple(range(1, len(dims) + 1))
    rng = np.random.default_rng(sum(dims) + batch)
    x = rng.standard_normal([batch] + dims)
    pairs = (x.reshape(-1, n)[:, 0::2] + 1j * x.reshape(-1, n)[:, 1::2]).reshape(-1)
    zq = emu.from_storage(emu.to_storage(pairs, fi), fi).reshape(-1, h)
    xq = np.empty((zq.shape[0], n))
    xq[:, 0::2], xq[:, 1::2] = zq.real, zq.imag
    xq = xq.reshape([batch] + dims)
    y = emu.real_exec_nd("r2c", x, dims, batch, fi, fo)
    ref = oracle.fft(xq.reshape(-1).astype(np.complex128), tuple(dims)).reshape([batch] + dims)[..., : h + 1]
    tol = TOL[fo if fo == emu.FMT_SPLIT else fi]
    assert oracle.rel_l2(y.reshape(-1), ref.reshape(-1)) <= tol
    spec = np.fft.rfftn(x, axes=axes) / np.sqrt(np.prod(dims))
    sq = emu.from_storage(emu.to_storage(spec.reshape(-1), fi), fi).reshape(spec.shape)
    back = emu.real_exec_nd("c2r", spec, dims, batch, fi, fo)
    refc = np.fft.irfftn(sq, s=dims, axes=axes) * np.prod(dims)
    assert oracle.rel_l2(back.reshape(-1).astype(np.complex128), refc.reshape(-1).astype(np.complex128)) <= tol


# ---- advanced layouts (strided / planar sides staged by gather and scatter) ----

def _dense(dims):
    s, st = 1, []
    for d in reversed(dims):
        st.append(s)
        s *= d
    return list(reversed(st)), s


LAYOUT_CASES = [
    # dims, batch, in layout (strides, dist, planar), out layout, in fmt, out fmt
    ([256], 3, ([3], 3 * 256 + 5, 0), None, emu.FMT_F16, emu.FMT_F16),                 # strided input only
    ([512], 2, None, ([2], 1100, 1), emu.FMT_BF16, emu.FMT_BF16),                      # planar output only
    ([64, 32], 2, ([1, 64], 64 * 32, 1), ([1, 64], 64 * 32, 1), emu.FMT_F16, emu.FMT_F16),  # column-major planar
    ([16, 16, 32], 1, ([16 * 40, 40, 1], 16 * 16 * 40, 0), ([16 * 40, 40, 1], 16 * 16 * 40, 1),
     emu.FMT_F16, emu.FMT_F16),                                                         # padded rows, 3D
    ([1024], 3, ([1], 1024, 1), ([5], 5 * 1024 + 1, 1), emu.FMT_F16, emu.FMT_SPLIT),  # planar FP16 -> planar split
    ([2048], 2, ([2], 4097, 0), ([1], 2048, 1), emu.FMT_SPLIT, emu.FMT_SPLIT),         # split, shared staging
]


@pytest.mark.parametrize("case", range(len(LAYOUT_CASES)))
@pytest.mark.parametrize("inverse", [False, True])
def test_layout_exec(case, inverse):
    """A layout exec equals the contiguous exec of the same values bit for bit
    (gather and scatter only move bits), matches the oracle, and leaves every
    output position outside the layout untouched."""
    dims, batch, li, lo, fi, fo = LAYOUT_CASES[case]
    st, n = _dense(dims)
    li = li or (st, n, 0)
    lo = lo or (st, n, 0)
    rng = np.random.default_rng(case)
    x = rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)
    y, (ore, oim), opos = emu.layout_exec(x, dims, batch, li, lo, fi, fo, inverse=inverse)
    y0, _ = emu.exec_fft(x, list(dims), batch, fi, fo, inverse=inverse)
    assert np.array_equal(y.reshape(-1), y0.reshape(-1)), (dims, li, lo)
    xq = emu.from_storage(emu.to_storage(x, fi), fi)
    ref = oracle.fft(xq, tuple(dims), sign=+1 if inverse else -1)
    assert oracle.rel_l2(y.reshape(-1), ref) <= TOL[fo if fo == emu.FMT_SPLIT else fi]
    # untouched positions keep the fill pattern
    for a in (ore, oim):
        if a is None:
            continue
        w = a.reshape(opos.max() + 1, -1)
        mask = np.ones(w.shape[0], dtype=bool)
        mask[opos] = False
        assert np.all(w[mask] == 0x7C01)


@pytest.mark.parametrize("case", [3, 4, 5])
def test_layout_four_element_path(case):
    """The four-element gather/scatter (innermost stride 1) moves the same bits
    as the one-element path."""
    dims, batch, li, lo, fi, fo = LAYOUT_CASES[case]
    st, n = _dense(dims)
    li = li or (st, n, 0)
    lo = lo or (st, n, 0)
    x = np.random.default_rng(11).standard_normal(n * batch) * (1 + 0.5j)
    y4, _, _ = emu.layout_exec(x, dims, batch, li, lo, fi, fo)
    assert emu.layout_vec_used()
    emu.layout_set_scalar_only(True)
    try:
        y1, _, _ = emu.layout_exec(x, dims, batch, li, lo, fi, fo)
        assert not emu.layout_vec_used()
    finally:
        emu.layout_set_scalar_only(False)
    assert np.array_equal(y4, y1)


def test_layout_planar_split_bits():
    """Planar split scalars are binary16 (hi, lo): the real-part array of a
    planar split output holds exactly the (re_hi, re_lo) halves of the
    interleaved split result."""
    dims, batch = [256], 2
    st, n = _dense(dims)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)
    _, (ore, oim), opos = emu.layout_exec(x, dims, batch, (st, n, 0), (st, n, 1), emu.FMT_F16, emu.FMT_SPLIT)
    _, (ire, _), _ = emu.layout_exec(x, dims, batch, (st, n, 0), (st, n + 1, 0), emu.FMT_F16, emu.FMT_SPLIT)
    inter = ire.reshape(-1, 4)
    pos = emu.layout_positions(dims, batch, st, n + 1)
    v = inter[pos]
    assert np.array_equal(ore.reshape(-1, 2), v[:, [0, 2]])
    assert np.array_equal(oim.reshape(-1, 2), v[:, [1, 3]])


def test_layout_invalid():
    """Zero strides or distance are invalid values; a planar side needs its
    imaginary array (checked by the exec's pointer rule)."""
    import ctypes
    L = emu.lib()
    dims = (ctypes.c_longlong * 1)(256)
    good = (ctypes.c_longlong * 3)(1, 0, 0)
    bad = (ctypes.c_longlong * 3)(0, 0, 0)
    buf = np.zeros(4096, dtype=np.uint16)
    e = ctypes.c_int()
    p = buf.ctypes.data
    emu.layout_exec(np.zeros(256), [256], 1, ([1], 256, 0), ([1], 256, 0), emu.FMT_F16, emu.FMT_F16)  # binds argtypes
    assert L.emu_layout_exec(1, dims, 1, bad, 256, 0, good, 256, 0, 0, 0, 1, 0, p, None, p, None, ctypes.byref(e)) == 1
    assert L.emu_layout_exec(1, dims, 1, good, 0, 0, good, 256, 0, 0, 0, 1, 0, p, None, p, None, ctypes.byref(e)) == 1
    assert L.emu_layout_exec(1, dims, 1, good, 256, 1, good, 256, 0, 0, 0, 1, 0, p, None, p, None, ctypes.byref(e)) == 1
    d3 = (ctypes.c_longlong * 1)(100)
    assert L.emu_layout_exec(1, d3, 1, good, 256, 0, good, 256, 0, 0, 0, 1, 0, p, None, p, None, ctypes.byref(e)) == 2


def test_column_views_fit_32bit_byte_offsets():
    """A column pass addresses (L - 1) * stride + Q elements through one 32-bit
    buffer view: 1024^3 fits in FP16 (4 GiB) but not in split (8 GiB), which
    the planner refuses instead of producing a truncated view."""
    assert "pass" in emu.describe([1024, 1024, 1024], 1, emu.FMT_F16, emu.FMT_F16)
    with pytest.raises(RuntimeError, match="status 2"):
        emu.describe([1024, 1024, 1024], 1, emu.FMT_F16, emu.FMT_SPLIT)
    assert "pass" in emu.describe([512, 1024, 1024], 1, emu.FMT_F16, emu.FMT_SPLIT)


def test_col_pass_byte_offsets_beyond_2gib():
    """The x-axis pass of an FP16 1024^3 plan addresses its columns through one
    buffer view of (1023 * 2^20 + Q) * 4 ~ 4.29e9 bytes, so half of every tile's
    element offsets become byte offsets above 2^31.  Run the tiles of that pass
    that cover columns 2^20 - 512 .. 2^20 - 1 of the y-z plane (every x row) in
    place on a 4 GiB anonymous mapping, of which only the touched pages are ever
    backed, and check each column against an FP64 FFT along x."""
    import mmap
    nx = 1 << 10
    plane = 1 << 20                      # ny * nz: the x-axis stride
    n = nx * plane
    flags = mmap.MAP_PRIVATE | mmap.MAP_ANONYMOUS | getattr(mmap, "MAP_NORESERVE", 0x4000)
    mm = mmap.mmap(-1, 4 * n, flags=flags)
    try:
        buf = np.frombuffer(mm, dtype=np.uint32)
        cols = np.arange(plane - 512, plane)
        idx = np.arange(nx)[:, None] * plane + cols[None, :]
        assert idx.max() * 4 > (1 << 31) and idx.max() * 4 < (1 << 32)
        rng = np.random.default_rng(1024)
        x = rng.standard_normal((nx, 512)) + 1j * rng.standard_normal((nx, 512))
        buf[idx] = emu.to_storage(x.reshape(-1), emu.FMT_F16).view(np.uint32).reshape(nx, 512)
        xq = emu.from_storage(buf[idx].reshape(-1), emu.FMT_F16).reshape(nx, 512)
        desc = emu.describe([nx, 1024, 1024], 1, emu.FMT_F16, emu.FMT_F16, 0)
        line = desc.splitlines()[3]
        assert "stride(in,out)=(1048576,1048576)" in line
        q = int(line.split()[1].split("_")[4])   # columns per tile of the x-axis pass
        ptr = ctypes.addressof(ctypes.c_char.from_buffer(mm))
        emu.exec_ptr(ptr, ptr, [nx, 1024, 1024], 1, emu.FMT_F16, emu.FMT_F16, scaling=0, only_pass=2,
                     blocks=(plane // q - 512 // q, 512 // q))
        y = emu.from_storage(buf[idx].reshape(-1), emu.FMT_F16).reshape(nx, 512)
        ref = np.fft.fft(xq, axis=0)
        err = oracle.rel_l2(y.reshape(-1), ref.reshape(-1))
        assert err <= 1.5e-3, err
        del buf, y
    finally:
        mm.close()


# ---------------- block floating point (MCFFT_SCALE_AUTO) ----------------
AUTO = emu.SCALE_AUTO


def sinusoid_input(dims, batch, seed, amplitude=100.0):
    """re/im iid U[-1,1) plus amplitude * exp(2 pi i (kx x / nx + ky y / ny)) at a
    random (kx, ky) per transform: the 2D benchmark input's shape (its spectrum
    peaks at amplitude * nx * ny)."""
    rng = np.random.default_rng(seed)
    nx, ny = dims
    xs, ys = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    out = []
    for _ in range(batch):
        kx, ky = rng.integers(0, nx), rng.integers(0, ny)
        u = rng.uniform(-1, 1, (nx, ny)) + 1j * rng.uniform(-1, 1, (nx, ny))
        out.append(u + amplitude * np.exp(2j * np.pi * (kx * xs / nx + ky * ys / ny)))
    return np.concatenate([o.reshape(-1) for o in out])


def small_uniform_input(n, seed):
    """re/im iid U[0, 1e-4): the magnitude of the reference's generate_data
    (source/mcfft_1d_utils.cpp:88-100), partly FP16 subnormals."""
    rng = np.random.default_rng(seed)
    return 1e-4 * (rng.uniform(0, 1, n) + 1j * rng.uniform(0, 1, n))


def run_auto(x, dims, batch, fi, fo, scaling=AUTO):
    xq = emu.from_storage(emu.to_storage(x, fi), fi)
    y, e = emu.exec_fft(x, list(dims), batch, fi, fo, scaling=scaling)
    return y, e, oracle.fft(xq, tuple(dims))


@pytest.mark.parametrize("dims,batch", [([256], 8), ([2048], 3), ([131072], 1), ([64, 128], 2), ([32, 32, 32], 1),
                                        ([16, 65536], 1)])
@pytest.mark.parametrize("fi,fo", FMTS)
def test_auto_scaling_parity(dims, batch, fi, fo):
    if fo == emu.FMT_SPLIT and dims[-1] > 16384:
        pytest.skip("split innermost sides stop at 2^14")
    check(dims, batch, fi, fo, scaling=AUTO)


@pytest.mark.parametrize("fi,fo", FMTS)
def test_auto_scaling_amplitude_100_sinusoid(fi, fo):
    """512 x 512: the peak 100 * 2^18 is far above FP16's 65504.  AUTO keeps the
    FP16/split result within tolerance with no status bit; BALANCED's static
    shifts (2^-8 over both axes) overflow FP16 and say so in the status word."""
    dims, batch = (512, 512), 2
    x = sinusoid_input(dims, batch, seed=3)
    y, e, ref = run_auto(x, dims, batch, fi, fo)
    err = oracle.rel_l2(y, ref)
    assert err <= TOL[fo if fo == emu.FMT_SPLIT else fi], err
    assert emu.status() == 0
    yb, _, _ = run_auto(x, dims, batch, fi, fo, scaling=emu.SCALE_BALANCED)
    if fi == emu.FMT_BF16:
        assert emu.status() == 0 and np.isfinite(yb).all()   # BF16's range holds it
    else:
        assert emu.status() == 2 and not np.isfinite(yb).all()


@pytest.mark.parametrize("dims,batch", [([4096], 2), ([256, 256], 1)])
def test_auto_scaling_small_inputs(dims, batch):
    """Inputs of generate_data's magnitude (~1e-4): AUTO scales them up and stays
    within the FP16 tolerance; FULL (divide by N) drops most of the spectrum
    below FP16's subnormal range."""
    n = int(np.prod(dims))
    x = small_uniform_input(n * batch, seed=42)
    y, e, ref = run_auto(x, dims, batch, emu.FMT_F16, emu.FMT_F16)
    assert oracle.rel_l2(y, ref) <= TOL[emu.FMT_F16]
    assert (e < 0).all()   # scaled up: stored values are larger than the true ones
    yf, _, _ = run_auto(x, dims, batch, emu.FMT_F16, emu.FMT_F16, scaling=emu.SCALE_FULL)
    assert oracle.rel_l2(yf, ref) > 0.05


@pytest.mark.parametrize("case", ["2d", "3d", "fourstep"])
@pytest.mark.parametrize("fi,fo", [(emu.FMT_F16, emu.FMT_F16), (emu.FMT_F16, emu.FMT_SPLIT)])
def test_auto_scaling_reconcile(case, fi, fo):
    """First-pass tiles of one transform with input maxima 2^4..2^12 apart pick
    different exponents; the pass after the first brings them to a common one
    element by element.  Parity against the oracle."""
    rng = np.random.default_rng(17)
    if case == "2d":
        dims, batch = (64, 256), 2
        scale = np.exp2(4.0 * (np.arange(64) % 4))[:, None] * np.ones((1, 256))        # per row
    elif case == "3d":
        dims, batch = (16, 32, 64), 1
        scale = np.exp2(4.0 * (np.arange(32) % 3))[None, :, None] * np.ones((16, 32, 64))  # per (x, y) row
    else:
        dims, batch = (131072,), 1
        scale = np.exp2(4.0 * ((np.arange(131072) % 512) // 32 % 3))              # per four-step column group
    n = int(np.prod(dims))
    x = (rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)) * np.tile(scale.reshape(-1), batch)
    y, e, ref = run_auto(x, dims, batch, fi, fo)
    assert oracle.rel_l2(y, ref) <= TOL[fo if fo == emu.FMT_SPLIT else fi]


def test_auto_scaling_zero_transforms():
    """An all-zero transform comes out as zeros; one whose rows are half zero,
    and a normal one next to it, keep their tolerance."""
    dims, batch = (64, 256), 3
    rng = np.random.default_rng(9)
    x = rng.standard_normal((batch, 64, 256)) + 1j * rng.standard_normal((batch, 64, 256))
    x[0] = 0
    x[1, ::2] = 0
    y, e, ref = run_auto(x.reshape(-1), dims, batch, emu.FMT_F16, emu.FMT_F16)
    y = y.reshape(batch, -1)
    ref = ref.reshape(batch, -1)
    assert (y[0] == 0).all()
    for b in (1, 2):
        assert oracle.rel_l2(y[b], ref[b]) <= TOL[emu.FMT_F16]


@pytest.mark.parametrize("dims", [[256], [64, 128]])
def test_auto_scaling_nonfinite_input(dims):
    """An Inf in one transform's input sets the status bit; the other
    transforms are unaffected."""
    batch = 4
    n = int(np.prod(dims))
    rng = np.random.default_rng(2)
    x = rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)
    x[2 * n + 5] = np.inf
    y, e = emu.exec_fft(x, dims, batch, emu.FMT_F16, emu.FMT_F16, scaling=AUTO)
    assert emu.status() == 2
    xq = emu.from_storage(emu.to_storage(x, emu.FMT_F16), emu.FMT_F16)
    y = y.reshape(batch, n)
    for b in (0, 1, 3):
        assert oracle.rel_l2(y[b], oracle.fft(xq[b * n:(b + 1) * n], tuple(dims))) <= TOL[emu.FMT_F16]
    x[2 * n + 5] = 0
    emu.exec_fft(x, dims, batch, emu.FMT_F16, emu.FMT_F16, scaling=AUTO)
    assert emu.status() == 0


@pytest.mark.parametrize("scaling", [AUTO, emu.SCALE_BALANCED])
@pytest.mark.parametrize("dims", [[1024], [131072], [32, 64]])
def test_input_exponents_add_to_output_exponents(scaling, dims):
    """Input exponents (true input = stored * 2^in[t]) add to the output
    exponents transform by transform, so the decoded results scale by 2^in[t]."""
    batch = 4 if dims != [131072] else 2
    n = int(np.prod(dims))
    rng = np.random.default_rng(4)
    x = rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)
    y0, e0 = emu.exec_fft(x, dims, batch, emu.FMT_F16, emu.FMT_F16, scaling=scaling)
    ins = np.array([3, -2, 0, 7][:batch], dtype=np.int32)
    emu.set_input_exponents(ins)
    try:
        y1, e1 = emu.exec_fft(x, dims, batch, emu.FMT_F16, emu.FMT_F16, scaling=scaling)
    finally:
        emu.set_input_exponents(None)
    assert (e1 == e0 + ins).all()
    assert np.array_equal(y1.reshape(batch, -1), y0.reshape(batch, -1) * np.exp2(ins.astype(float))[:, None])


@pytest.mark.parametrize("kernel", ["mcfft_pass_bigauto16_0", "mcfft_pass_bigauto_0"])
def test_auto_scaling_single_pass_65536(kernel):
    """The AUTO form of the single-pass 65536 kernel (normalises and writes the
    exponent of each transform itself), both wave counts, with transforms of
    very different magnitude side by side."""
    emu.set_kernel_override(kernel)
    try:
        n, batch = 65536, 3
        rng = np.random.default_rng(12)
        x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))) * np.array([1e-4, 1.0, 300.0])[:, None]
        y, e, ref = run_auto(x.reshape(-1), [n], batch, emu.FMT_F16, emu.FMT_F16)
        assert kernel in emu.last_kernels()
        y, ref = y.reshape(batch, n), ref.reshape(batch, n)
        for b in range(batch):
            assert oracle.rel_l2(y[b], ref[b]) <= TOL[emu.FMT_F16]
        assert e[0] < e[1] < e[2]
    finally:
        emu.set_kernel_override(None)
