"""mcfft_error_norms_noalloc (partials in library-owned device memory) against
mcfft_error_norms (partials in a buffer allocated per call), through the C ABI.

Both run the same reduction kernel with the same grid and combine the partials
in the same order on the host, so the results must be bit-identical.  Two calls
in a row, the second on a smaller input, check that nothing left in the shared
partials array by the first call leaks into the second.
"""
import numpy as np
import pytest

import mcfft_amd as mc

pytestmark = pytest.mark.gpu


def _case(torch, n, fmt, seed, exponent=0):
    """y = x in storage format `fmt`; ref = (x + 1e-3 noise) * 2^exponent, so
    y * 2^exponent is close to ref."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = (x + 1e-3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))) * 2.0 ** exponent
    return mc.to_device(x, fmt), torch.from_numpy(ref.view(np.float64).copy()).cuda()


@pytest.mark.parametrize("fmt", ["fp16", "bf16", "split"])
def test_noalloc_matches_allocating_entry_point(torch_cuda, fmt):
    torch = torch_cuda
    for n, seed, exponent in ((1 << 20, 5, 2), (4097, 6, -1)):   # second call smaller
        y, ref = _case(torch, n, fmt, seed, exponent)
        r0, m0 = mc.error_norms(y, fmt, exponent, ref)
        r1, m1 = mc.error_norms(y, fmt, exponent, ref, noalloc=True)
        print(f"{fmt} n={n}: rel_l2 {r0:.6e} / {r1:.6e}, max_abs {m0:.6e} / {m1:.6e}")
        assert r1 == r0 and m1 == m0
        assert 0.0 < r0 < 0.05   # quantisation of x plus the 1e-3 noise


def test_noalloc_rejects_invalid_sizes(torch_cuda):
    """Sizes outside 1 .. 2^40 elements are refused before any GPU work."""
    import ctypes
    torch = torch_cuda
    y, ref = _case(torch, 64, "fp16", 7)
    r, m = ctypes.c_double(), ctypes.c_double()
    for n in (0, -5, (1 << 40) + 1):
        st = mc.lib().mcfft_error_norms_noalloc(y.data_ptr(), 0, 0, ref.data_ptr(), n, ctypes.byref(r),
                                                ctypes.byref(m), None)
        assert st == 1   # MCFFT_INVALID_VALUE
