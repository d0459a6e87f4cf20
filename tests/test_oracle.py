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


def test_round_half_known_values():
    v = np.array([1.0, 1 + 2 ** -11, 1 + 3 * 2 ** -11, 65504.0, 65520.0, 2 ** -25, 3 * 2 ** -25])
    r = oracle.round_half(v.astype(complex)).real
    assert r[0] == 1.0 and r[1] == 1.0            # tie to even
    assert r[2] == 1 + 2 * 2 ** -10               # tie to even (up)
    assert r[3] == 65504.0 and np.isinf(r[4])
    assert r[5] == 0.0 and r[6] == 2 ** -23       # subnormal ties (1.5 ulp -> 2 ulp)
    assert np.array_equal(r, v.astype(np.float16).astype(np.float64))
