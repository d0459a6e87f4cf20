"""Plan structures on the CPU emulation of the kernels (tests/emu), against the
FP64 oracle -- CPU only.

* 2D plans with a split outer axis: NX = R * NX/R, a row pass (ROWX) that takes
  the R-point DFT across rows NX/R apart with twiddle W_NX^(r1*k2) before its
  row FFTs, then NX/R-point columns.  Every compiled ROWX variant, forward and
  inverse, in every format; AUTO scaling with first-pass tiles that disagree
  (the column pass reconciles them); in-place execs and chunked batches.
* The structures the autotuner enumerates (outer radix, four-step split), plans
  built with each, and the versioned tuning-cache format."""
import ctypes

import numpy as np
import pytest

import emu
import oracle

TOL = {emu.FMT_F16: 1.5e-3, emu.FMT_BF16: 1.2e-2, emu.FMT_SPLIT: 1e-6}
FMTS = [(emu.FMT_F16, emu.FMT_F16), (emu.FMT_BF16, emu.FMT_BF16), (emu.FMT_F16, emu.FMT_SPLIT)]
AUTO = 3


def kernels():
    return emu.last_kernels().split()


def run(dims, batch, fi, fo, x, inverse=False, scaling=AUTO, ws_limit=0):
    y, e = emu.exec_fft(x, list(dims), batch, fi, fo, scaling=scaling, inverse=inverse, ws_limit=ws_limit)
    xq = emu.from_storage(emu.to_storage(x, fi), fi)
    ref = oracle.fft(xq, tuple(dims), sign=+1 if inverse else -1)
    return oracle.rel_l2(y, ref), e


# every compiled ROWX variant: row length L (512..4096), radix R (the tile's rows)
ROWX = [(L, R) for L in (512, 1024, 2048, 4096) for R in (2, 4)] + [(2048, 8), (4096, 8)]


@pytest.mark.parametrize("L,R", ROWX)
@pytest.mark.parametrize("fi,fo", FMTS)
def test_rowx_variant(L, R, fi, fo, monkeypatch):
    """(16 R) x L with the outer axis split by R: the ROWX kernel with R rows per
    tile, then 16-point columns; forward for R = 2, 8 and inverse for R = 4.
    BF16 and split run every variant up to 2048 points and 4096 x 2 (the other
    4096-point variants share their code paths with the FP16 ones)."""
    monkeypatch.setenv("MCFFT_2D_OUTER_RADIX", str(R))
    if fi != emu.FMT_F16 or fo != emu.FMT_F16:
        if L == 4096 and R > 2:
            pytest.skip("BF16/split 4096-point ROWX: 4096 x 2 covers the row length (CPU time)")
    dims = (16 * R, L)
    rng = np.random.default_rng(L + R)
    n = dims[0] * dims[1]
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    err, _ = run(dims, 1, fi, fo, x, inverse=(R == 4))
    assert kernels()[0].endswith("_%d_5" % fi if fo != emu.FMT_SPLIT else "_2_5"), kernels()
    assert f"_{R}_" in kernels()[0], kernels()
    assert err <= TOL[fo if fo == emu.FMT_SPLIT else fi], err


@pytest.mark.parametrize("fi,fo", FMTS)
@pytest.mark.parametrize("scaling", [0, 1, 2, AUTO])
def test_rowx_scalings(fi, fo, scaling, monkeypatch):
    """128 x 512 split by 4, every scaling mode, batch 2."""
    monkeypatch.setenv("MCFFT_2D_OUTER_RADIX", "4")
    rng = np.random.default_rng(7)
    n = 128 * 512 * 2
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    err, e = run((128, 512), 2, fi, fo, x, scaling=scaling)
    assert "_4_1_" in kernels()[0] and kernels()[0].endswith("_5"), kernels()
    assert err <= TOL[fo if fo == emu.FMT_SPLIT else fi], err
    if scaling == 2:
        assert list(e) == [16, 16]   # FULL: every stage divides by its radix, log2(2^16)


@pytest.mark.parametrize("fi,fo", FMTS)
def test_rowx_auto_reconcile(fi, fo, monkeypatch):
    """AUTO with first-pass tiles 2^8 apart in magnitude: rows r1 + 64 r2 of
    ROWX tile r1 are scaled by 2^8 for r1 < 32 only, so the tiles choose
    different exponents and the column pass brings them to a common one.  Same
    error bound as unscaled data, relative to the transform's norm."""
    monkeypatch.setenv("MCFFT_2D_OUTER_RADIX", "4")
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, (256, 512)) + 1j * rng.uniform(-1, 1, (256, 512))
    r = np.arange(256) % 64
    x = x * np.where(r < 32, 256.0, 1.0)[:, None]
    err, e = run((256, 512), 1, fi, fo, x.reshape(-1))
    assert err <= TOL[fo if fo == emu.FMT_SPLIT else fi], err


def test_rowx_in_place_and_chunks(monkeypatch):
    """In place (input = output buffer: the row pass goes through the plan's
    workspace) and a batch run as one transform per chunk."""
    monkeypatch.setenv("MCFFT_2D_OUTER_RADIX", "2")
    dims, batch = (128, 512), 3
    n = dims[0] * dims[1]
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)
    buf = emu.to_storage(x, emu.FMT_F16).copy()
    xq = emu.from_storage(buf.copy(), emu.FMT_F16)
    e = emu.exec_ptr(buf.ctypes.data, buf.ctypes.data, list(dims), batch, emu.FMT_F16, emu.FMT_F16, scaling=1)
    y = emu.from_storage(buf, emu.FMT_F16) * 2.0 ** e
    ref = oracle.fft(xq, dims)
    assert oracle.rel_l2(y, ref) <= TOL[emu.FMT_F16]
    # one transform per chunk: 4 bytes per element
    err, _ = run(dims, batch, emu.FMT_F16, emu.FMT_F16, x, ws_limit=n * 4)
    assert err <= TOL[emu.FMT_F16]


def test_rowx_default_structure(monkeypatch):
    """Which 2D plans split their outer axis, and by what."""
    monkeypatch.delenv("MCFFT_2D_OUTER_RADIX", raising=False)
    F16, SPL = emu.FMT_F16, emu.FMT_SPLIT
    cases = [((4096, 4096), F16, F16, 2), ((4096, 4096), F16, SPL, 2), ((2048, 2048), F16, SPL, 4),
             ((1024, 1024), F16, F16, 4), ((1024, 4096), F16, F16, 4), ((512, 512), F16, F16, 0),
             ((4096, 256), F16, F16, 0)]
    for dims, fi, fo, r in cases:
        d = emu.describe(list(dims), 1, fi, fo, AUTO)
        assert f"struct=R={r};" in d, (dims, d)
        first = d.splitlines()[1].split()[1]
        assert first.endswith("_5") == (r > 0), (dims, d)


def test_structures_enumerated():
    """The autotuner's candidates: every outer radix a 2D shape can take (one
    entry per distinct plan), the four-step split and its neighbours in 1D, and
    nothing for shapes with a single structure."""
    F16, SPL = emu.FMT_F16, emu.FMT_SPLIT
    assert emu.structures([4096, 4096], 4, F16, F16, AUTO) == ["R=2;L2=0", "R=0;L2=0", "R=4;L2=0", "R=8;L2=0"]
    # no split ROWX kernel holds 8 rows of 4096 points: R = 8 plans as R = 0
    assert emu.structures([4096, 4096], 4, F16, SPL, AUTO) == ["R=2;L2=0", "R=0;L2=0", "R=4;L2=0"]
    assert emu.structures([1 << 20], 2, F16, F16, AUTO) == ["R=0;L2=10", "R=0;L2=9", "R=0;L2=11"]
    assert emu.structures([1 << 23], 1, F16, F16, AUTO) == ["R=0;L2=11"]   # N1 <= 2^12, N2 <= 2^11
    # single-pass lengths from 2^12 up may also run as two passes
    assert emu.structures([4096], 8, F16, F16, AUTO) == ["R=0;L2=0", "R=0;L2=5", "R=0;L2=6", "R=0;L2=7"]
    assert emu.structures([2048], 8, F16, F16, AUTO) == ["R=0;L2=0"]
    assert emu.structures([64, 64, 64], 1, F16, F16, AUTO) == ["R=0;L2=0"]
    d = emu.describe_opts([1 << 20], 2, F16, F16, AUTO, fourstep_lg2=9)
    assert "struct=R=0;L2=9" in d and "stride(in,out)=(512,512)" in d, d


@pytest.mark.parametrize("n,lg2", [(1 << 17, 8), (1 << 17, 10), (1 << 18, 10), (1 << 13, 6), (1 << 16, 8)])
@pytest.mark.parametrize("fi,fo", FMTS)
def test_fourstep_split_choices(n, lg2, fi, fo):
    """Two-pass 1D plans with a non-default N1 x N2 split (as the autotuner may
    choose) against the oracle."""
    rng = np.random.default_rng(n + lg2)
    x = rng.standard_normal(2 * n) + 1j * rng.standard_normal(2 * n)
    emu.set_structure(-1, lg2)
    try:
        err, _ = run([n], 2, fi, fo, x, inverse=(lg2 == 10))
        d = emu.describe_opts([n], 2, fi, fo, AUTO, fourstep_lg2=lg2)
    finally:
        emu.set_structure(-1, -1)
    assert f"L2={lg2}" in d
    assert err <= TOL[fo if fo == emu.FMT_SPLIT else fi], err


def test_tune_cache_versioning():
    """A tuning cache is read only by the build that wrote it, in its format."""
    key = "gfx950|2|4096x4096|4|0|0|3"
    val = "R=2;L2=0;K=mcfft_pass_3_1_2_8_0_5,mcfft_pass_2_8_16_16_0_2"
    text = "mcfft-tune-cache 2 abc123\n" + key + "=" + val + "\n"
    assert emu.tune_cache_lookup(text, "abc123", key) == val
    assert emu.tune_cache_lookup(text, "def456", key) is None          # another build
    assert emu.tune_cache_lookup(text.replace("cache 2", "cache 1"), "abc123", key) is None  # another format
    assert emu.tune_cache_lookup(text, "abc123", key + "x") is None
    assert emu.tune_cache_lookup("mcfft-tune-cache 2 abc123\n" + key + "=K=a,b\n", "abc123", key) is None


# ---------------- gathered batches (a ragged plan's same-shape items) ----------------
GATHER_CASES = [([4096], 5, 3), ([1 << 17], 3, 0), ([128, 512], 3, 0), ([32, 32, 32], 2, 3), ([8192], 3, 1)]


@pytest.mark.parametrize("dims,count,scaling", GATHER_CASES)
@pytest.mark.parametrize("fi,fo", [(emu.FMT_F16, emu.FMT_F16), (emu.FMT_F16, emu.FMT_SPLIT)])
def test_gathered_batch(dims, count, scaling, fi, fo, monkeypatch):
    """Transforms at scattered offsets (gaps, reversed order, input and output
    placed differently) through one batched plan that locates them with offset
    tables (the 2D case with its outer axis split: ROWX through the workspace):
    each against the oracle, and nothing outside the transforms written."""
    monkeypatch.setenv("MCFFT_2D_OUTER_RADIX", "4")
    if scaling == AUTO:
        scaling = 1   # ragged plans are statically scaled
    n = int(np.prod(dims))
    rng = np.random.default_rng(n + count)
    slot_in = rng.permutation(count + 2)[:count]          # slots of n (+ a gap) elements
    slot_out = rng.permutation(count + 1)[:count]
    in_offs = slot_in * (n + 37)
    out_offs = slot_out[::-1] * (n + 5)
    x = rng.standard_normal((count, n)) + 1j * rng.standard_normal((count, n))
    comp_in = 4 if fi == emu.FMT_SPLIT else 2
    comp_out = 4 if fo == emu.FMT_SPLIT else 2
    buf_in = np.zeros(((count + 2) * (n + 37)) * comp_in, dtype=np.uint16)
    for t in range(count):
        buf_in[in_offs[t] * comp_in:(in_offs[t] + n) * comp_in] = emu.to_storage(x[t], fi)
    sentinel = 0x7BCD
    buf_out = np.full(((count + 1) * (n + 5)) * comp_out, sentinel, dtype=np.uint16)
    e = emu.exec_gathered(buf_in, buf_out, dims, in_offs, out_offs, fi, fo, scaling=scaling)
    written = np.zeros(buf_out.size, dtype=bool)
    for t in range(count):
        sl = slice(out_offs[t] * comp_out, (out_offs[t] + n) * comp_out)
        written[sl] = True
        y = emu.from_storage(buf_out[sl], fo) * 2.0 ** e[t]
        xq = emu.from_storage(emu.to_storage(x[t], fi), fi)
        err = oracle.rel_l2(y, oracle.fft(xq, tuple(dims)))
        assert err <= TOL[fo if fo == emu.FMT_SPLIT else fi], (t, err)
    assert np.all(buf_out[~written] == sentinel)


@pytest.mark.parametrize("n", [2048, 4096, 8192, 16384, 32768])
def test_one_row_groups_take_one_row_tiles(n):
    """A row pass with one row per group (a single transform here; every
    transform of a gathered batch) runs a kernel whose tile holds one row, where
    one is compiled, instead of a tile of several rows run mostly empty."""
    for fmt in (emu.FMT_F16, emu.FMT_BF16):
        d = emu.describe([n], 1, fmt, fmt)
        name = [l.split()[1] for l in d.splitlines() if l.strip().startswith("pass")][0]
        assert name.split("_")[4] == "1", d


@pytest.mark.parametrize("name", ["mcfft_pass_3_1_2_4_0_5", "mcfft_pass_3_1_2_4_1_5", "mcfft_pass_3_1_2_4_2_5"])
def test_rowx_two_column_sets_per_wave(name, monkeypatch):
    """ROWX tiles of 4096-point rows on 4 waves (each wave takes two 32-column
    sets of every row and pairs the rows' sets for its butterfly): 32 x 4096
    forward and inverse against the oracle."""
    monkeypatch.setenv("MCFFT_2D_OUTER_RADIX", "2")
    cm = int(name.split("_")[-2])
    fi, fo = (emu.FMT_F16, emu.FMT_SPLIT) if cm == 2 else (cm, cm)
    rng = np.random.default_rng(len(name))
    x = rng.uniform(-1, 1, 32 * 4096) + 1j * rng.uniform(-1, 1, 32 * 4096)
    emu.set_kernel_override(name)
    try:
        for inverse in (False, True):
            err, _ = run((32, 4096), 1, fi, fo, x, inverse=inverse)
            assert kernels()[0] == name, kernels()
            assert err <= TOL[fo if fo == emu.FMT_SPLIT else fi], err
    finally:
        emu.set_kernel_override(None)


# ---------------- 2D/3D rows longer than one pass ----------------
LONG_ROWS = [((16, 32768), emu.FMT_F16, emu.FMT_SPLIT, AUTO)]


@pytest.mark.parametrize("dims,fi,fo,scaling", LONG_ROWS)
def test_long_rows_two_pass(dims, fi, fo, scaling):
    """Innermost sides above the single-pass limit (split 2^15): a two-pass
    four-step over every row, then the column pass; AUTO with rows 2^8 apart in
    magnitude (first-pass tiles that disagree).  The GPU tests run the larger
    shapes (split 2^16 rows, FP16/BF16 2^17 rows, 3D)."""
    n = int(np.prod(dims))
    rng = np.random.default_rng(n)
    x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).reshape(-1, dims[-1])
    x[::2] *= 256.0
    d = emu.describe(list(dims), 1, fi, fo, scaling)
    assert "fourstep" in d and len(d.splitlines()) == len(dims) + 2, d
    err, _ = run(dims, 1, fi, fo, x.reshape(-1), scaling=scaling, inverse=(scaling == 1))
    assert err <= TOL[fo if fo == emu.FMT_SPLIT else fi], err
