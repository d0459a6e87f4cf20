"""Pins for the CPU oracle (oracle/oracle.c) -- CPU only.

Each check is chosen so that a plausible mistake (a sign, an index, a dropped
term, a transposed operand) fails at least one of them:
  * the reference's own captured outputs (tests/golden, from 3.txt / 2.txt / 1.log)
  * brute-force long-double DFT
  * closed forms: delta, constant, single exponential, shift/modulation theorems,
    Parseval, linearity, separability of 2D/3D, inverse round trip
  * numpy.fft (an independent implementation)
"""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")
# the captured values are printed with 7 significant digits (%e)
PRINT_TOL = 3e-7


def test_generate_data_and_fft_match_reference_3txt():
    g = np.load(os.path.join(GOLD, "ref_3txt_n256_b256_seed42.npz"))
    x = oracle.generate_data(int(g["n"]), int(g["n_batch"]), int(g["seed"]))
    y = oracle.fft(x, (256,))
    assert oracle.rel_l2(y[:2048], g["head"]) < PRINT_TOL
    assert oracle.rel_l2(y[g["sample_idx"]], g["sample_val"]) < PRINT_TOL


def test_fft_matches_reference_fftw_2txt():
    g = np.load(os.path.join(GOLD, "ref_2txt_n262144_b1_seed42.npz"))
    x = oracle.generate_data(262144, 1, 42)
    y = oracle.fft(x, (262144,))
    assert oracle.rel_l2(y[:2048], g["fftw_head"]) < PRINT_TOL
    assert oracle.rel_l2(y[g["fftw_sample_idx"]], g["fftw_sample_val"]) < PRINT_TOL
    # the reference's own column of that dump is wrong (kept as evidence)
    assert oracle.rel_l2(g["mcfft_head"], y[:2048]) > 0.1


def test_gen_rev_matches_1log():
    g = np.load(os.path.join(GOLD, "ref_1log_genrev512.npz"))
    assert np.array_equal(oracle.gen_rev(512, [16, 16, 2]), g["rev"])
    # one radix -> identity (the hook the compatibility plan relies on)
    assert np.array_equal(oracle.gen_rev(4096, [4096]), np.arange(4096))


def test_planted_sign_error_is_caught():
    """A forward/inverse sign mix-up must fail the golden check."""
    g = np.load(os.path.join(GOLD, "ref_3txt_n256_b256_seed42.npz"))
    x = oracle.generate_data(256, 8, 42)
    wrong = oracle.fft(x, (256,), sign=+1)
    assert oracle.rel_l2(wrong[:2048], g["head"]) > 1e-2


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 16, 30, 64, 97, 256, 1000, 1024])
def test_against_bruteforce(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert oracle.rel_l2(oracle.fft(x, (n,)), oracle.dft_naive(x)) < 1e-14
    assert oracle.rel_l2(oracle.fft(x, (n,), sign=+1), oracle.dft_naive(x, sign=+1)) < 1e-14


@pytest.mark.parametrize("n", [16, 256, 4096, 65536])
def test_closed_forms(n):
    d = np.zeros(n, complex)
    d[0] = 1
    assert np.allclose(oracle.fft(d, (n,)), np.ones(n), atol=1e-15)
    ones = np.ones(n, complex)
    e = np.zeros(n, complex)
    e[0] = n
    assert np.allclose(oracle.fft(ones, (n,)), e, atol=1e-9 * n)
    k0 = 3 % n
    ex = np.exp(2j * np.pi * k0 * np.arange(n) / n)
    t = np.zeros(n, complex)
    t[k0] = n
    assert np.max(np.abs(oracle.fft(ex, (n,)) - t)) < 1e-10 * n
    # shift theorem: x[n-s] -> X[k] W^(ks)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    s = 5
    X = oracle.fft(x, (n,))
    Xs = oracle.fft(np.roll(x, s), (n,))
    assert oracle.rel_l2(Xs, X * np.exp(-2j * np.pi * np.arange(n) * s / n)) < 1e-13
    # Parseval and linearity
    assert abs(np.sum(np.abs(X) ** 2) / n - np.sum(np.abs(x) ** 2)) / np.sum(np.abs(x) ** 2) < 1e-13
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert oracle.rel_l2(oracle.fft(2 * x - 3j * y, (n,)), 2 * X - 3j * oracle.fft(y, (n,))) < 1e-13
    # inverse round trip (unnormalised)
    assert oracle.rel_l2(oracle.fft(X, (n,), sign=+1) / n, x) < 1e-14


@pytest.mark.parametrize("dims", [(8, 16), (64, 32), (4, 8, 16), (16, 16, 16)])
def test_nd_separable_and_numpy(dims):
    rng = np.random.default_rng(2)
    b = 2
    x = rng.standard_normal((b,) + dims) + 1j * rng.standard_normal((b,) + dims)
    y = oracle.fft(x, dims).reshape((b,) + dims)
    z = x.copy()
    for ax in range(1, len(dims) + 1):  # 1D oracle along each axis
        z = np.moveaxis(z, ax, -1)
        sh = z.shape
        z = oracle.fft(z.reshape(-1), (sh[-1],)).reshape(sh)
        z = np.moveaxis(z, -1, ax)
    assert oracle.rel_l2(y, z) < 1e-14
    assert oracle.rel_l2(y, np.fft.fftn(x, axes=tuple(range(1, len(dims) + 1)))) < 1e-14


@pytest.mark.parametrize("lg,batch", [(2, 3), (9, 2), (15, 1), (20, 2)])
def test_fft_1d_large_matches_stockham(lg, batch):
    """The four-step path (single long transforms, parallel inside one
    transform) equals the pinned Stockham path, both directions."""
    n = 1 << lg
    rng = np.random.default_rng(lg)
    x = rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)
    for sign in (-1, 1):
        a = oracle.fft_large(x, n, sign)
        b = x.copy()
        oracle.lib().orc_fft_many(oracle._dptr(b.view(np.float64)), 1, (oracle.ctypes.c_long * 1)(n), batch, sign)
        assert oracle.rel_l2(a, b) < 1e-14, sign


def test_reference_metric_definition():
    s = np.array([1 + 1j, 2 - 2j])
    t = np.array([1.5 + 1j, 2 - 10j])
    # |0.5/1| + 0 + 0 + min(1, 8/2) = 1.5 over 2 complex
    assert abs(oracle.ref_metric(t, s) - 0.75) < 1e-15


@pytest.mark.parametrize("n", [256, 512, 1024, 131072])
def test_reference_fp16_emulator_algebra(n):
    """With rounding disabled the emulated reference kernels are the exact FFT."""
    b = 1 if n > 4096 else 4
    x = oracle.generate_data(n, b, 42)
    e = oracle.ref_fp16_exec(x, n, rounding=False)
    assert oracle.rel_l2(e, oracle.fft(x, (n,))) < 1e-14


@pytest.mark.parametrize("n", [256, 512])
def test_reference_fp16_emulator_error_band(n):
    """At FP16 the reference algorithm's error is O(2^-11) per stage."""
    rng = np.random.default_rng(4)
    x = rng.standard_normal(n * 8) + 1j * rng.standard_normal(n * 8)
    xq = oracle.round_half(x)
    err = oracle.rel_l2(oracle.ref_fp16_exec(xq, n), oracle.fft(xq, (n,)))
    assert 1e-4 < err < 5e-3


# ---- the FP16 emulator's rounding points, pinned by what the mathematics fixes ----
U = 2.0 ** -11   # FP16 unit roundoff
# radix-16 MMA stages per length (the 512/1024 tails are VALU radix-2/4 steps)
MMA_STAGES = {256: 2, 512: 2, 1024: 2, 131072: 4}


def _f16(a):
    """numpy's own FP16 rounding (independent of oracle.round_half)."""
    return np.asarray(a).astype(np.float16).astype(np.float64)


def _gauss_rho2():
    """Relative variance of FP16 rounding of a Gaussian (scale-invariant under
    powers of two, and within 0.5% across an octave)."""
    v = np.random.default_rng(0).standard_normal(2_000_000)
    return float(np.mean((_f16(v) - v) ** 2) / np.mean(v ** 2))


@pytest.mark.parametrize("n", [256, 512, 1024, 131072])
@pytest.mark.parametrize("fused", [False, True])
def test_reference_fp16_emulator_exact_inputs(n, fused):
    """FP16-exact inputs.  A delta gives exactly all ones at every rounding
    point: every value it reaches is 1 or 0 (column 0 of every table and every
    twiddle it meets is 1).  A single exponential at k0 in {0, N/4, N/2}
    (entries 0, +-1, +-i) gives exactly N in the real part of bin k0; the other
    bins hold only what the tables' zeros leave, at the 2^-24 level: rounded as
    the reference builds them (cosf of a float argument, so cos(pi/2) becomes
    -2^-24 in FP16), or FP64 tables whose products the MFMA takes in FP32.
    131072 exceeds FP16's range (65504), so its exponentials are not run."""
    extra = oracle.R_FUSED if fused else 0
    d = np.zeros(n, complex)
    d[0] = 1
    assert np.array_equal(oracle.ref_fp16_exec(d, n, points=oracle.R_ALL | extra), np.ones(n))
    if n > 65504:
        return
    m = np.arange(n)
    for k0 in (0, n // 4, n // 2):
        x = np.exp(2j * np.pi * k0 * m / n)
        x = np.round(x.real) + 1j * np.round(x.imag)
        want = np.zeros(n, complex)
        want[k0] = n
        for pts in (oracle.R_ALL, oracle.R_ALL & ~oracle.R_TABLE):
            y = oracle.ref_fp16_exec(x, n, points=pts | extra)
            assert y[k0].real == n, (k0, pts)
            assert np.max(np.abs(y - want)) <= n * 2.0 ** -18, (k0, pts)


@pytest.mark.parametrize("n", [256, 512, 1024, 131072])
def test_reference_fp16_emulator_mma_error_model(n):
    """First-order model of the half accumulator: every radix-16 stage rounds
    each output once and, per mma_sync rounding, also the partial -(Fi Xi) --
    half the output's variance on the 14 rows whose Fi is not zero -- so
    rel_l2^2 = rho^2 S (1 + 7/16), or rho^2 S when the four mma_sync share one
    FP32 accumulator (rho^2: FP16 rounding variance of a Gaussian, S: stages).
    Band +-2%; the two predictions are 20% apart, so moving the rounding point
    fails the check (asserted both ways)."""
    rho2 = _gauss_rho2()
    b = max(1, (1 << 19) // n)
    rng = np.random.default_rng(1)
    xq = oracle.round_half(rng.standard_normal(n * b) + 1j * rng.standard_normal(n * b))
    ref = oracle.fft(xq, (n,))
    s = MMA_STAGES[n]
    pred = {False: np.sqrt(rho2 * s * (1 + 7 / 16)), True: np.sqrt(rho2 * s)}
    for fused in (False, True):
        pts = oracle.R_MMA | (oracle.R_FUSED if fused else 0)
        meas = oracle.rel_l2(oracle.ref_fp16_exec(xq, n, points=pts), ref)
        assert abs(meas / pred[fused] - 1) < 0.02, (fused, meas / U, pred[fused] / U)
        assert abs(meas / pred[not fused] - 1) > 0.08, fused


def _tw_chain(N, col, row, tables, valu):
    """The stated twiddle recurrence: t = W(N, col*4*(row//4)), then
    row % 4 times t = t * W(N, col); rounded tables and/or rounded complex
    products (two products and their sum/difference each rounded)."""
    def w(k):
        z = np.exp(-2j * np.pi * k / N)
        return complex(_f16(z.real), _f16(z.imag)) if tables else z
    t, f = w(col * 4 * (row // 4)), w(col)
    for _ in range(row % 4):
        if valu:
            t = complex(_f16(_f16(t.real * f.real) - _f16(t.imag * f.imag)),
                        _f16(_f16(t.real * f.imag) + _f16(t.imag * f.real)))
        else:
            t = t * f
    return t


def _factorised_256(x, tables, valu, chain=True):
    """X[k1 + 16 k2] = sum_m2 F[k2, m2] T[k1, m2] sum_m1 F[k1, m1] x[16 m1 + m2]
    (Cooley-Tukey 16 x 16) with the 16-point matrix F and the twiddles T rounded
    to FP16 (tables), and the twiddle products rounded (valu); chain=False takes
    T straight from the rounded table (a moved rounding point)."""
    k = np.arange(16)
    F = np.exp(-2j * np.pi * np.outer(k, k) / 16)
    if tables:
        F = _f16(F.real) + 1j * _f16(F.imag)
    if chain:
        T = np.array([[_tw_chain(256, k1, m2, tables, valu) for m2 in k] for k1 in k])
    else:
        T = np.exp(-2j * np.pi * np.outer(k, k) / 256)
        T = _f16(T.real) + 1j * _f16(T.imag)
    X = x.reshape(-1, 16, 16)
    Y = np.einsum("km,bmn->bkn", F, X)
    P = Y * T[None]
    if valu:
        tr, ti = T.real[None], T.imag[None]
        P = (_f16(_f16(Y.real * tr) - _f16(Y.imag * ti))
             + 1j * _f16(_f16(Y.real * ti) + _f16(Y.imag * tr)))
    return np.einsum("jn,bkn->bjk", F, P).reshape(-1)


def test_reference_fp16_emulator_table_and_twiddle_points_256():
    """The emulator's table and twiddle-arithmetic rounding points against the
    textbook 16 x 16 factorisation with the same stated roundings (numpy's own
    FP16 conversion, FP64 tables): the error each class adds agrees to 1e-4 (the
    only differences are ulp-level flips where a float-argument cosf and an
    FP64 cos round apart), and taking the twiddles straight from the table
    instead of the recurrence moves the error by more than 10%.  With every
    point on, the classes add in quadrature with the MMA model within 2%."""
    rng = np.random.default_rng(2)
    n, b = 256, 2048
    xq = oracle.round_half(rng.standard_normal(n * b) + 1j * rng.standard_normal(n * b))
    ref = oracle.fft(xq, (n,))
    meas = {}
    for tables, valu in ((True, False), (False, True), (True, True)):
        pts = (oracle.R_TABLE if tables else 0) | (oracle.R_VALU if valu else 0)
        meas[tables, valu] = oracle.rel_l2(oracle.ref_fp16_exec(xq, n, points=pts), ref)
        model = oracle.rel_l2(_factorised_256(xq, tables, valu), ref)
        assert abs(meas[tables, valu] / model - 1) < 1e-4, (tables, valu)
        moved = oracle.rel_l2(_factorised_256(xq, tables, valu, chain=False), ref)
        assert abs(meas[tables, valu] / moved - 1) > 0.10, (tables, valu)
    rho2 = _gauss_rho2()
    for fused in (False, True):
        pts = oracle.R_ALL | (oracle.R_FUSED if fused else 0)
        whole = oracle.rel_l2(oracle.ref_fp16_exec(xq, n, points=pts), ref)
        mma2 = rho2 * MMA_STAGES[n] * (1 if fused else 1 + 7 / 16)
        assert abs(whole / np.sqrt(meas[True, True] ** 2 + mma2) - 1) < 0.02, fused


@pytest.mark.parametrize("n", [256, 512, 1024, 131072])
def test_reference_fp16_emulator_variants_bracket(n):
    """The fused-accumulator variant rounds a subset of the default's points,
    so on random inputs it is the more accurate (the tighter bound)."""
    rng = np.random.default_rng(3)
    b = max(1, 4096 // n)
    xq = oracle.round_half(rng.standard_normal(n * b) + 1j * rng.standard_normal(n * b))
    ref = oracle.fft(xq, (n,))
    per_mma = oracle.rel_l2(oracle.ref_fp16_exec(xq, n), ref)
    fused = oracle.rel_l2(oracle.ref_fp16_exec(xq, n, points=oracle.R_ALL | oracle.R_FUSED), ref)
    assert fused < per_mma


def test_round_half_known_values():
    v = np.array([1.0, 1 + 2 ** -11, 1 + 3 * 2 ** -11, 65504.0, 65520.0, 2 ** -25, 3 * 2 ** -25])
    r = oracle.round_half(v.astype(complex)).real
    assert r[0] == 1.0 and r[1] == 1.0            # tie to even
    assert r[2] == 1 + 2 * 2 ** -10               # tie to even (up)
    assert r[3] == 65504.0 and np.isinf(r[4])
    assert r[5] == 0.0 and r[6] == 2 ** -23       # subnormal ties (1.5 ulp -> 2 ulp)
    assert np.array_equal(r, v.astype(np.float16).astype(np.float64))
