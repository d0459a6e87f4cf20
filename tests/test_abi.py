"""The C ABI boundary -- CPU only: libmcfft.so loads and exports every function
include/mcfft.h declares, the reference-compatible shim libraries export the
reference's C++ signatures, and argument errors are reported as status codes
before any GPU work (no compute call is made here)."""
import ctypes
import os
import re
import subprocess

import pytest

import mcfft_amd as mc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared_functions(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mcfft_[a-z0-9_]+)\s*\(", text)))


def dynamic_symbols(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", lib], capture_output=True, text=True, check=True).stdout
    return out


def test_library_exports_every_declared_function():
    names = declared_functions("mcfft.h")
    assert len(names) >= 12
    L = mc.lib()
    for n in names:
        assert hasattr(L, n), n
    syms = dynamic_symbols(mc.LIB_PATH)
    for n in names:
        assert re.search(rf"\bT {n}\b", syms), n


@pytest.mark.parametrize("dim,expected", [
    (1, ["mcfftCreate(mcfftHandle*, int, int)", "mcfftExec(mcfftHandle, __half*)", "setup(double*, int, int)",
         "doit(int)", "finalize(double*)", "generate_data(double*, int, int, int)"]),
    (2, ["mcfftCreate(mcfftHandle*, int, int, int)", "mcfftExec(mcfftHandle, __half*)",
         "setup(double*, int, int, int)", "doit(int)", "finalize(double*)", "generate_data(double*, int, int, int)"]),
    (3, ["mcfftCreate(mcfftHandle*, int, int, int)", "mcfftCreate(mcfftHandle*, int, int, int, int)",
         "setup(double*, int, int, int)", "setup(double*, int, int, int, int)", "doit(int)"]),
])
def test_compat_libraries_export_reference_signatures(dim, expected):
    syms = dynamic_symbols(os.path.join(mc.LIB_DIR, f"libmcfft_compat_{dim}d.so"))
    for e in expected:
        assert e in syms, e


def test_status_strings_and_version():
    L = mc.lib()
    for s in range(8):
        assert L.mcfft_status_string(s)
    assert L.mcfft_version() >= 100


def test_argument_errors_are_status_codes():
    L = mc.lib()
    h = ctypes.c_void_p()
    dims = (ctypes.c_longlong * 1)(256)
    assert L.mcfft_plan_many(None, 1, dims, 1, 0, 0, 1) == 1          # null plan pointer
    assert L.mcfft_plan_many(ctypes.byref(h), 0, dims, 1, 0, 0, 1) == 1   # rank 0
    assert L.mcfft_plan_many(ctypes.byref(h), 1, dims, 0, 0, 0, 1) == 1   # batch 0
    assert L.mcfft_plan_many(ctypes.byref(h), 1, dims, 1, 0, 0, 7) == 1   # scaling enum
    assert L.mcfft_plan_many(ctypes.byref(h), 1, dims, 1, 2, 0, 1) == 2   # split -> fp16
    bad = (ctypes.c_longlong * 1)(1000)
    assert L.mcfft_plan_many(ctypes.byref(h), 1, bad, 1, 0, 0, 1) == 2   # not a power of two
    # outside the documented sizes: 1D below 16 or above 2^27, an outer side above
    # 4096, fewer columns on a 16-point outer axis than the 32 of its column tile
    for shape in [(8,), (1 << 28,), (8192, 64), (16, 16), (16, 16, 16)]:
        d = (ctypes.c_longlong * len(shape))(*shape)
        assert L.mcfft_plan_many(ctypes.byref(h), len(shape), d, 1, 0, 0, 1) == 2, shape
    assert L.mcfft_exec(None, None, None, -1, None) == 6                  # invalid plan
    assert L.mcfft_destroy(None) == 6
    e = ctypes.c_int()
    assert L.mcfft_get_output_exponent(None, ctypes.byref(e)) == 6


CALLER = r"""
#include "mcfft_%dd_utils.h"
// a caller of the reference's type surface: benchmark/test_1d.cpp:105 declares
// vkfft_get_result(MCTYPE2* data, MCTYPE* result, int n, int n_batch)
static void pairs_to_scalars(MCTYPE2* data, MCTYPE* result, int n, int n_batch) {
  for (long long i = 0; i < (long long)n * n_batch; ++i) {
    result[2 * i] = data[i].x;
    result[2 * i + 1] = data[i].y;
  }
}
int main() {
  static_assert(sizeof(MCTYPE2) == 2 * sizeof(MCTYPE), "pair of scalars");
  MCTYPE2 d[4];
  MCTYPE r[8];
  pairs_to_scalars(d, r, 2, 2);
  double buf[32];
  generate_data(buf, 16, 1);
  return 0;
}
"""


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("bf16", [False, True])
def test_compat_headers_declare_reference_types(tmp_path, dim, bf16):
    """The shim headers define MCTYPE and MCTYPE2 (include/common.h:4-5 of the
    reference), so reference callers written against them compile."""
    src = tmp_path / "caller.cpp"
    src.write_text(CALLER % dim)
    cmd = ["/opt/rocm/bin/hipcc", "-fsyntax-only", "-std=c++17", f"-I{INC}", str(src)]
    if bf16:
        cmd.insert(1, "-DMCFFT_COMPAT_BF16")
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
