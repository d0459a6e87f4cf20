"""CPU oracle for the batched C2C FFT path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The shipped
library (``mcfft_amd``) never imports it and has no CPU fallback.

The numerical work lives in ``oracle.c`` (compiled by ``oracle/Makefile`` into
``oracle/_build/liboracle.so``); this module is a thin numpy/ctypes wrapper.
Reference citations for every routine are in ``oracle.c``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile oracle.c with gcc (called by __graft_entry__.build())."""
    if force or not os.path.exists(_LIB_PATH) or (
        os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, "oracle.c"))
    ):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_fft_many.argtypes = [dp, ctypes.c_int, ctypes.POINTER(ctypes.c_long), ctypes.c_long, ctypes.c_int]
        L.orc_fft_many.restype = None
        L.orc_fft_1d_large.argtypes = [dp, ctypes.c_long, ctypes.c_long, ctypes.c_int]
        L.orc_fft_1d_large.restype = None
        L.orc_dft_naive.argtypes = [dp, dp, ctypes.c_long, ctypes.c_int]
        L.orc_dft_naive.restype = None
        L.orc_generate_data.argtypes = [dp, ctypes.c_long, ctypes.c_long, ctypes.c_int]
        L.orc_generate_data.restype = None
        L.orc_gen_rev.argtypes = [ctypes.c_long, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.orc_gen_rev.restype = None
        for name in ("orc_ref_metric", "orc_rel_l2", "orc_max_abs"):
            getattr(L, name).argtypes = [dp, dp, ctypes.c_long]
            getattr(L, name).restype = ctypes.c_double
        L.orc_ref_fp16_exec.argtypes = [dp, dp, ctypes.c_long, ctypes.c_long, ctypes.c_int]
        L.orc_ref_fp16_exec.restype = ctypes.c_int
        L.orc_round_half.argtypes = [dp, ctypes.c_long]
        L.orc_round_half.restype = None
        L.orc_num_threads.argtypes = []
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _as_c128(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.complex128))


def fft(x, axes_shape, sign: int = -1) -> np.ndarray:
    """FP64 batched FFT.  ``x`` is complex (batch, *axes_shape) or flat; the last
    axis is contiguous.  sign=-1 forward, +1 inverse, both unnormalised."""
    a = _as_c128(x).copy()
    dims = (ctypes.c_long * len(axes_shape))(*axes_shape)
    per = int(np.prod(axes_shape))
    batch = a.size // per
    if len(axes_shape) == 1 and per >= LARGE_1D and batch < num_threads() and per & (per - 1) == 0:
        lib().orc_fft_1d_large(_dptr(a.view(np.float64)), per, batch, int(sign))
    else:
        lib().orc_fft_many(_dptr(a.view(np.float64)), len(axes_shape), dims, batch, int(sign))
    return a


# single 1D transforms from this length on (fewer than the threads) run the
# four-step path, parallel inside each transform
LARGE_1D = 1 << 20


def fft_large(x, n: int, sign: int = -1) -> np.ndarray:
    """The four-step path directly (for its pin against orc_fft_many)."""
    a = _as_c128(x).copy()
    lib().orc_fft_1d_large(_dptr(a.view(np.float64)), n, a.size // n, int(sign))
    return a


def dft_naive(x, sign: int = -1) -> np.ndarray:
    a = _as_c128(x)
    out = np.empty_like(a)
    lib().orc_dft_naive(_dptr(a.view(np.float64)), _dptr(out.view(np.float64)), a.size, int(sign))
    return out


def generate_data(n: int, n_batch: int, seed: int = 42) -> np.ndarray:
    """Restatement of the reference generator; returns complex128 (n_batch*n,)."""
    out = np.empty(2 * n * n_batch, dtype=np.float64)
    lib().orc_generate_data(_dptr(out), n, n_batch, seed)
    return out.view(np.complex128)


def gen_rev(n: int, radices) -> np.ndarray:
    r = (ctypes.c_int * len(radices))(*radices)
    out = np.empty(n, dtype=np.int32)
    lib().orc_gen_rev(n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), r, len(radices))
    return out


def rel_l2(a, ref) -> float:
    a = _as_c128(a)
    ref = _as_c128(ref)
    return lib().orc_rel_l2(_dptr(a.view(np.float64)), _dptr(ref.view(np.float64)), a.size)


def max_abs(a, ref) -> float:
    a = _as_c128(a)
    ref = _as_c128(ref)
    return lib().orc_max_abs(_dptr(a.view(np.float64)), _dptr(ref.view(np.float64)), a.size)


def ref_metric(tested, standard) -> float:
    t = _as_c128(tested)
    s = _as_c128(standard)
    return lib().orc_ref_metric(_dptr(t.view(np.float64)), _dptr(s.view(np.float64)), t.size)


# rounding points of ref_fp16_exec (flags; see oracle.c)
R_INPUT, R_TABLE, R_MMA, R_VALU, R_FUSED = 1, 2, 4, 8, 16
R_ALL = R_INPUT | R_TABLE | R_MMA | R_VALU


def ref_fp16_exec(x, n: int, rounding: bool = True, points: int | None = None) -> np.ndarray:
    """Emulate the reference kernels at MCTYPE=half on natural-order input.

    rounding=False keeps everything in FP64; ``points`` (R_* flags) selects the
    rounding points instead: R_ALL is the default, R_ALL | R_FUSED the variant
    whose complex_mul keeps FP32 across its four mma_sync."""
    a = _as_c128(x)
    out = np.empty_like(a)
    mode = (points << 1) if points is not None else int(bool(rounding))
    if points is not None and points == 0:
        mode = 0
    rc = lib().orc_ref_fp16_exec(_dptr(a.view(np.float64)), _dptr(out.view(np.float64)), n, a.size // n, mode)
    if rc != 0:
        raise ValueError(f"the reference has no plan for n={n}")
    return out


def round_half(x) -> np.ndarray:
    a = _as_c128(x).copy()
    lib().orc_round_half(_dptr(a.view(np.float64)), 2 * a.size)
    return a


def num_threads() -> int:
    return lib().orc_num_threads()
