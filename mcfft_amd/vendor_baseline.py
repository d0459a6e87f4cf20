"""rocFFT (the ROCm FFT library) on the same buffers, as the GPU vendor baseline
beside this library's numbers (the reference benchmarks against rocFFT in
benchmark/test_1d.cpp:30-103 and benchmark/test_2d.cpp).

Only measurement uses this module; the library itself never calls rocFFT.

    p = RocfftPlan([131072], 256, "half")     # or "single"
    p.exec(in_ptr, out_ptr, stream)           # interleaved complex, out of place
"""
from __future__ import annotations

import ctypes
import os

_lib = None

PRECISIONS = {"single": 0, "double": 1, "half": 2}
ELEM_BYTES = {"single": 8, "double": 16, "half": 4}


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "librocfft.so")
        L = ctypes.CDLL(path)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.rocfft_setup.restype = ctypes.c_int
        L.rocfft_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, sz,
                                         ctypes.POINTER(sz), sz, vp]
        L.rocfft_plan_create.restype = ctypes.c_int
        L.rocfft_plan_get_work_buffer_size.argtypes = [vp, ctypes.POINTER(sz)]
        L.rocfft_plan_get_work_buffer_size.restype = ctypes.c_int
        L.rocfft_execution_info_create.argtypes = [ctypes.POINTER(vp)]
        L.rocfft_execution_info_create.restype = ctypes.c_int
        L.rocfft_execution_info_set_work_buffer.argtypes = [vp, vp, sz]
        L.rocfft_execution_info_set_work_buffer.restype = ctypes.c_int
        L.rocfft_execution_info_set_stream.argtypes = [vp, vp]
        L.rocfft_execution_info_set_stream.restype = ctypes.c_int
        L.rocfft_execute.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), vp]
        L.rocfft_execute.restype = ctypes.c_int
        L.rocfft_plan_destroy.argtypes = [vp]
        L.rocfft_execution_info_destroy.argtypes = [vp]
        if L.rocfft_setup() != 0:
            raise RuntimeError("rocfft_setup failed")
        _lib = L
    return _lib


def _check(st, what):
    if st != 0:
        raise RuntimeError(f"{what}: rocfft status {st}")


class RocfftPlan:
    """Batched C2C plan (row-major dims, last contiguous), out of place."""

    def __init__(self, dims, batch: int, precision: str = "half", inverse: bool = False, real: bool = False):
        """real: real-input forward (inverse=False) or real-output inverse transform."""
        import torch
        L = lib()
        lengths = (ctypes.c_size_t * len(dims))(*reversed([int(d) for d in dims]))  # fastest first
        self._plan = ctypes.c_void_p()
        ttype = (3 if inverse else 2) if real else (1 if inverse else 0)
        _check(L.rocfft_plan_create(ctypes.byref(self._plan), 1, ttype, PRECISIONS[precision],
                                    len(dims), lengths, int(batch), None), "rocfft_plan_create")
        ws = ctypes.c_size_t()
        _check(L.rocfft_plan_get_work_buffer_size(self._plan, ctypes.byref(ws)), "rocfft_plan_get_work_buffer_size")
        self._info = ctypes.c_void_p()
        _check(L.rocfft_execution_info_create(ctypes.byref(self._info)), "rocfft_execution_info_create")
        self._work = None
        if ws.value:
            self._work = torch.empty(ws.value, dtype=torch.uint8, device="cuda")
            _check(L.rocfft_execution_info_set_work_buffer(self._info, self._work.data_ptr(), ws.value),
                   "rocfft_execution_info_set_work_buffer")
        self.work_bytes = ws.value

    def exec(self, in_ptr: int, out_ptr: int, stream: int):
        L = lib()
        _check(L.rocfft_execution_info_set_stream(self._info, stream), "rocfft_execution_info_set_stream")
        i = (ctypes.c_void_p * 1)(in_ptr)
        o = (ctypes.c_void_p * 1)(out_ptr)
        _check(L.rocfft_execute(self._plan, i, o, self._info), "rocfft_execute")

    def destroy(self):
        L = lib()
        if self._plan:
            L.rocfft_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()
        if self._info:
            L.rocfft_execution_info_destroy(self._info)
            self._info = ctypes.c_void_p()
