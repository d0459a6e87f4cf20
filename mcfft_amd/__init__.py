"""Python host binding of the MI355X batched C2C FFT library (include/mcfft.h).

The compute path is ``mcfft_amd/_lib/libmcfft.so`` (hand-written gfx950 HIP
kernels behind a C ABI); this module only wraps plan/exec/destroy with ctypes
and converts between complex arrays and the device storage formats.  There is
no CPU fallback: if the shared library is missing or HIP fails, calls raise.

Storage formats as torch tensors (last axis = components of one element):
  "fp16"   torch.float16  (..., 2)  re, im
  "bf16"   torch.bfloat16 (..., 2)  re, im
  "split"  torch.float16  (..., 4)  re_hi, im_hi, re_lo, im_lo; value = hi + lo * 2^-11
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libmcfft.so")
# MCFFT_LIB: load another build of the same library (e.g. an experimental variant)
LIB_PATH = os.environ.get("MCFFT_LIB", LIB_PATH)

FORMATS = {"fp16": 0, "bf16": 1, "split": 2}
SCALINGS = {"none": 0, "balanced": 1, "full": 2, "auto": 3}
STATUS_WAIT_BOUND, STATUS_NONFINITE = 1, 2
FORWARD, INVERSE = -1, 1
STATUS = {0: "success", 1: "invalid value", 2: "unsupported", 3: "allocation failed", 4: "HIP error",
          5: "misaligned buffer", 6: "invalid plan", 7: "tuning cache error", 8: "non-finite result (overflow)"}

_lib = None


class RaggedItem(ctypes.Structure):
    _fields_ = [("rank", ctypes.c_int), ("dims", ctypes.c_longlong * 3), ("in_offset", ctypes.c_longlong),
                ("out_offset", ctypes.c_longlong)]


class Layout(ctypes.Structure):
    """mcfftLayout: element (b, i_0, ..) at b*dist + sum_d i_d*strides[d]."""
    _fields_ = [("strides", ctypes.c_longlong * 3), ("dist", ctypes.c_longlong), ("planar", ctypes.c_int)]

    @classmethod
    def make(cls, strides, dist: int, planar: bool = False) -> "Layout":
        v = cls()
        for j, s in enumerate(strides):
            v.strides[j] = int(s)
        v.dist, v.planar = int(dist), int(bool(planar))
        return v


class McfftError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: {STATUS.get(status, status)} (status {status})")
        self.status = status


def lib() -> ctypes.CDLL:
    """Load libmcfft.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, ll, i = ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int
        L.mcfft_plan_many.argtypes = [ctypes.POINTER(vp), i, ctypes.POINTER(ll), ll, i, i, i]
        L.mcfft_plan_1d.argtypes = [ctypes.POINTER(vp), ll, ll, i, i]
        L.mcfft_plan_2d.argtypes = [ctypes.POINTER(vp), ll, ll, ll, i, i]
        L.mcfft_plan_3d.argtypes = [ctypes.POINTER(vp), ll, ll, ll, ll, i, i]
        L.mcfft_exec.argtypes = [vp, vp, vp, i, vp]
        L.mcfft_exec_exp.argtypes = [vp, vp, vp, vp, i, vp]
        L.mcfft_get_transform_exponents.argtypes = [vp, ctypes.POINTER(i)]
        L.mcfft_get_exponent_buffer.argtypes = [vp, ctypes.POINTER(vp)]
        L.mcfft_exec_passes.argtypes = [vp, vp, vp, i, vp, i, i]
        L.mcfft_exec_timed.argtypes = [vp, vp, vp, i, vp, ctypes.POINTER(vp), ctypes.POINTER(vp)]
        L.mcfft_get_pass_count.argtypes = [vp, ctypes.POINTER(i)]
        L.mcfft_get_output_exponent.argtypes = [vp, ctypes.POINTER(i)]
        L.mcfft_get_workspace_size.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
        L.mcfft_plan_describe.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
        L.mcfft_destroy.argtypes = [vp]
        L.mcfft_status_string.argtypes = [i]
        L.mcfft_status_string.restype = ctypes.c_char_p
        L.mcfft_version.restype = i
        L.mcfft_plan_ragged.argtypes = [ctypes.POINTER(vp), i, ctypes.POINTER(RaggedItem), i, i, i]
        L.mcfft_plan_many_layout.argtypes = [ctypes.POINTER(vp), i, ctypes.POINTER(ll), ll, ctypes.POINTER(Layout),
                                             ctypes.POINTER(Layout), i, i, i]
        L.mcfft_exec_layout.argtypes = [vp, vp, vp, vp, vp, i, vp]
        L.mcfft_plan_r2c_1d.argtypes = [ctypes.POINTER(vp), ll, ll, i, i, i]
        L.mcfft_plan_c2r_1d.argtypes = [ctypes.POINTER(vp), ll, ll, i, i, i]
        L.mcfft_plan_r2c.argtypes = [ctypes.POINTER(vp), i, ctypes.POINTER(ll), ll, i, i, i]
        L.mcfft_plan_c2r.argtypes = [ctypes.POINTER(vp), i, ctypes.POINTER(ll), ll, i, i, i]
        L.mcfft_get_item_exponents.argtypes = [vp, ctypes.POINTER(i)]
        L.mcfft_get_status_flags.argtypes = [vp, ctypes.POINTER(ctypes.c_uint)]
        L.mcfft_plan_tune.argtypes = [vp, vp, vp, i, vp, i]
        L.mcfft_error_norms.argtypes = [vp, i, i, vp, ll, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double), vp]
        L.mcfft_error_norms_noalloc.argtypes = L.mcfft_error_norms.argtypes
        L.mcfft_build_id.argtypes = [ctypes.POINTER(ctypes.c_char_p)]
        for f in ("mcfft_plan_many", "mcfft_plan_1d", "mcfft_plan_2d", "mcfft_plan_3d", "mcfft_exec",
                  "mcfft_exec_exp", "mcfft_get_transform_exponents", "mcfft_get_exponent_buffer",
                  "mcfft_exec_passes", "mcfft_exec_timed", "mcfft_get_pass_count", "mcfft_plan_ragged", "mcfft_get_item_exponents",
                  "mcfft_get_status_flags", "mcfft_plan_tune", "mcfft_error_norms", "mcfft_error_norms_noalloc",
                  "mcfft_build_id",
                  "mcfft_plan_r2c_1d",
                  "mcfft_plan_c2r_1d", "mcfft_plan_r2c", "mcfft_plan_c2r", "mcfft_plan_many_layout", "mcfft_exec_layout",
                  "mcfft_get_output_exponent", "mcfft_get_workspace_size", "mcfft_plan_describe", "mcfft_destroy"):
            getattr(L, f).restype = i
        _lib = L
    return _lib


def build_id() -> str:
    """Identifier of the kernel sources libmcfft.so was built from (the tuning
    cache records it and ignores caches of other builds)."""
    v = ctypes.c_char_p()
    _check(lib().mcfft_build_id(ctypes.byref(v)), "mcfft_build_id")
    return v.value.decode()


def _check(st: int, what: str):
    if st != 0:
        raise McfftError(st, what)


ELEM_BYTES = {"fp16": 4, "bf16": 4, "split": 8}


def _check_tensor(t, nbytes: int, what: str):
    """The C ABI takes bare pointers: a device tensor smaller than what the plan
    reads or writes would be overrun by the kernels, so check it here."""
    if t is None:
        raise ValueError(f"{what}: missing tensor")
    if not t.is_cuda:
        raise ValueError(f"{what}: not a device tensor")
    if not t.is_contiguous():
        raise ValueError(f"{what}: not contiguous")
    have = t.numel() * t.element_size()
    if have < nbytes:
        raise ValueError(f"{what}: {have} bytes, the plan needs {nbytes}")


def _static_exponent(h, scaling: str):
    """The plan's one exponent (static scaling), or None (AUTO: per transform)."""
    if scaling == "auto":
        return None
    e = ctypes.c_int()
    _check(lib().mcfft_get_output_exponent(h, ctypes.byref(e)), "mcfft_get_output_exponent")
    return e.value


class Plan:
    """A plan for `batch` transforms of shape `dims` (1 to 3 dims, last contiguous).

    scaling "auto" (default): block floating point following the data, one
    exponent per transform found on the GPU (`exponents()`, `result()`);
    "none" / "balanced" / "full": one static exponent (`exponent`)."""

    def __init__(self, dims, batch: int, in_fmt: str = "fp16", out_fmt: str | None = None,
                 scaling: str = "auto"):
        dims = [int(d) for d in (dims if hasattr(dims, "__len__") else [dims])]
        out_fmt = out_fmt or in_fmt
        self.dims, self.batch, self.in_fmt, self.out_fmt, self.scaling = dims, int(batch), in_fmt, out_fmt, scaling
        self._h = ctypes.c_void_p()
        arr = (ctypes.c_longlong * len(dims))(*dims)
        _check(lib().mcfft_plan_many(ctypes.byref(self._h), len(dims), arr, self.batch, FORMATS[in_fmt],
                                     FORMATS[out_fmt], SCALINGS[scaling]), "mcfft_plan_many")
        self.exponent = _static_exponent(self._h, scaling)
        self.n = int(np.prod(dims))

    def exponents(self) -> np.ndarray:
        """Output exponent of every transform of the last exec (true result =
        stored * 2^e[t]); waits for the device."""
        e = (ctypes.c_int * self.batch)()
        _check(lib().mcfft_get_transform_exponents(self._h, e), "mcfft_get_transform_exponents")
        return np.array(e, dtype=np.int64)

    def exponent_ptr(self) -> int:
        """Device address of the plan's per-transform exponent array (batch
        int32), written by its execs in stream order (usable as another exec's
        in_exp)."""
        p = ctypes.c_void_p()
        _check(lib().mcfft_get_exponent_buffer(self._h, ctypes.byref(p)), "mcfft_get_exponent_buffer")
        return p.value

    def result(self, out, count: int | None = None) -> np.ndarray:
        """The first `count` transforms (default: all) of device output `out` of
        the last exec, decoded in true units with their own exponents."""
        count = self.batch if count is None else int(count)
        per = self.output_elems_per_transform()
        y = from_device(out[: count * per], self.out_fmt)
        return (y.reshape(count, per) * np.exp2(self.exponents()[:count].astype(np.float64))[:, None]).reshape(-1)

    def output_elems_per_transform(self) -> int:
        return self.n

    @property
    def handle(self) -> int:
        return self._h.value

    def workspace_bytes(self) -> int:
        b = ctypes.c_size_t()
        _check(lib().mcfft_get_workspace_size(self._h, ctypes.byref(b)), "mcfft_get_workspace_size")
        return b.value

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(16384)
        _check(lib().mcfft_plan_describe(self._h, buf, 16384), "mcfft_plan_describe")
        return buf.value.decode()

    def exec_ptr(self, in_ptr: int, out_ptr: int, direction: int = FORWARD, stream: int = 0, in_exp: int = 0):
        """mcfft_exec_exp on raw device pointers (no size checks); in_exp: device
        address of per-transform input exponents, or 0."""
        _check(lib().mcfft_exec_exp(self._h, in_ptr, in_exp or None, out_ptr, direction, stream), "mcfft_exec_exp")

    def io_bytes(self):
        """(input bytes, output bytes) of the contiguous transform: what exec
        reads and writes, and what tune / exec_passes / exec_timed use."""
        return self.n * self.batch * ELEM_BYTES[self.in_fmt], self.n * self.batch * ELEM_BYTES[self.out_fmt]

    def exec_io_bytes(self):
        """(input bytes, output bytes) that exec reads and writes."""
        return self.io_bytes()

    def _check_io(self, inp, out, contiguous=False):
        ib, ob = self.io_bytes() if contiguous else self.exec_io_bytes()
        _check_tensor(inp, ib, "input")
        _check_tensor(out, ob, "output")

    def exec(self, inp, out, direction: int = FORWARD, stream=None, in_exp=None):
        """Run on torch device tensors (storage layout above) on `stream`
        (default: torch's current stream).  in_exp: the input's per-transform
        exponents -- a Plan whose last exec produced `inp` (its exponent array),
        a device int32 tensor of batch entries, or None (all 0)."""
        import torch
        self._check_io(inp, out)
        if stream is None:
            stream = torch.cuda.current_stream(inp.device).cuda_stream
        elif hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        ptr = 0
        if isinstance(in_exp, Plan):
            if in_exp.batch != self.batch:
                raise ValueError("in_exp plan has another batch")
            ptr = in_exp.exponent_ptr()
        elif in_exp is not None:
            _check_tensor(in_exp, 4 * self.batch, "in_exp")
            if in_exp.dtype != torch.int32:
                raise ValueError("in_exp must be int32")
            ptr = in_exp.data_ptr()
        self.exec_ptr(inp.data_ptr(), out.data_ptr(), direction, stream, ptr)

    def tune(self, inp, out, direction: int = FORWARD, reps: int = 5, stream=None):
        """Autotune on these buffers: for each plan structure (2D outer-axis radix,
        1D four-step split) every compiled kernel variant of each pass, then whole
        execs; keep the fastest and record it in the tuning cache (blocking;
        overwrites out, may replace the plan's device buffers)."""
        import torch
        self._check_io(inp, out, contiguous=True)
        if stream is None:
            stream = torch.cuda.current_stream(inp.device).cuda_stream
        _check(lib().mcfft_plan_tune(self._h, inp.data_ptr(), out.data_ptr(), direction, stream, reps),
               "mcfft_plan_tune")

    def pass_count(self) -> int:
        c = ctypes.c_int()
        _check(lib().mcfft_get_pass_count(self._h, ctypes.byref(c)), "mcfft_get_pass_count")
        return c.value

    def status_flags(self) -> int:
        """Status bits (STATUS_*; 0 = fine; waits for the device): regular plans
        report an Inf/NaN stored (static scaling) or read (AUTO) by any exec
        since the last call, and clear it; ragged plans a dependency wait of
        the last exec that hit its bound."""
        f = ctypes.c_uint()
        _check(lib().mcfft_get_status_flags(self._h, ctypes.byref(f)), "mcfft_get_status_flags")
        return f.value

    def exec_passes(self, inp, out, direction: int, first: int, count: int, stream=None):
        """Launch only passes [first, first+count) (profiling; leaves `out` incomplete)."""
        import torch
        self._check_io(inp, out, contiguous=True)
        if stream is None:
            stream = torch.cuda.current_stream(inp.device).cuda_stream
        _check(lib().mcfft_exec_passes(self._h, inp.data_ptr(), out.data_ptr(), direction, stream, first, count),
               "mcfft_exec_passes")

    def exec_timed(self, inp, out, direction, start_events, stop_events, stream=None):
        """exec with each pass's kernel bracketed on the GPU by its own pair of
        torch.cuda.Events (already created, e.g. recorded once): elapsed_time
        between start_events[i] and stop_events[i] is pass i's kernel duration."""
        import torch
        self._check_io(inp, out, contiguous=True)
        if stream is None:
            stream = torch.cuda.current_stream(inp.device).cuda_stream
        n = self.pass_count()
        if len(start_events) != n or len(stop_events) != n:
            raise ValueError(f"exec_timed needs {n} start and {n} stop events (one pair per pass), got "
                             f"{len(start_events)} and {len(stop_events)}")
        st = (ctypes.c_void_p * n)(*[e.cuda_event for e in start_events])
        sp = (ctypes.c_void_p * n)(*[e.cuda_event for e in stop_events])
        _check(lib().mcfft_exec_timed(self._h, inp.data_ptr(), out.data_ptr(), direction, stream, st, sp),
               "mcfft_exec_timed")

    def destroy(self):
        if self._h and self._h.value:
            lib().mcfft_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class RaggedPlan(Plan):
    """One persistent launch over many transforms of different shapes.
    `items`: list of (dims, in_offset, out_offset) with element offsets."""

    def __init__(self, items, in_fmt: str = "fp16", out_fmt: str | None = None, scaling: str = "balanced"):
        # ragged items are statically scaled ("auto" is refused)
        out_fmt = out_fmt or in_fmt
        self.in_fmt, self.out_fmt, self.scaling = in_fmt, out_fmt, scaling
        self.items = [(list(d), int(io), int(oo)) for d, io, oo in items]
        arr = (RaggedItem * len(items))()
        for k, (d, io, oo) in enumerate(self.items):
            arr[k].rank = len(d)
            for j in range(3):
                arr[k].dims[j] = d[j] if j < len(d) else 1
            arr[k].in_offset, arr[k].out_offset = io, oo
        self._h = ctypes.c_void_p()
        _check(lib().mcfft_plan_ragged(ctypes.byref(self._h), len(items), arr, FORMATS[in_fmt], FORMATS[out_fmt],
                                       SCALINGS[scaling]), "mcfft_plan_ragged")
        ex = (ctypes.c_int * len(items))()
        _check(lib().mcfft_get_item_exponents(self._h, ex), "mcfft_get_item_exponents")
        self.exponents = list(ex)
        self.exponent = None
        # extents checked on every exec: computed once here (an O(items) loop per
        # exec would leave the GPU idle for milliseconds at tens of thousands of items)
        ie = oe = 0
        for d, io, oo in self.items:
            n = 1
            for v in d:
                n *= int(v)
            ie, oe = max(ie, io + n), max(oe, oo + n)
        self._io_bytes = (ie * ELEM_BYTES[in_fmt], oe * ELEM_BYTES[out_fmt])

    def io_bytes(self):
        """Bytes up to the end of the furthest item, in and out."""
        return self._io_bytes


class RealPlan(Plan):
    """Real-input forward ("r2c": dims reals -> outer x (n/2+1) complex per
    transform, exec with FORWARD) or real-output inverse ("c2r": outer x (n/2+1)
    complex -> dims reals, exec with INVERSE); `dims` is n (1D) or the list of 1
    to 3 dimensions, n the last.  Real data is stored as complex pairs
    (x[2m], x[2m+1]) along the last dimension."""

    def __init__(self, dims, batch: int, kind: str = "r2c", in_fmt: str = "fp16", out_fmt: str | None = None,
                 scaling: str = "auto"):
        dims = [int(d) for d in (dims if hasattr(dims, "__len__") else [dims])]
        out_fmt = out_fmt or in_fmt
        self.dims, self.batch, self.in_fmt, self.out_fmt, self.scaling = dims, int(batch), in_fmt, out_fmt, scaling
        self.kind = kind
        self._h = ctypes.c_void_p()
        arr = (ctypes.c_longlong * len(dims))(*dims)
        fn = lib().mcfft_plan_r2c if kind == "r2c" else lib().mcfft_plan_c2r
        _check(fn(ctypes.byref(self._h), len(dims), arr, int(batch), FORMATS[in_fmt], FORMATS[out_fmt],
                  SCALINGS[scaling]), f"mcfft_plan_{kind}")
        self.exponent = _static_exponent(self._h, scaling)
        self.n = dims[-1]
        self.rows = int(np.prod(dims[:-1])) if len(dims) > 1 else 1

    def output_elems_per_transform(self) -> int:
        """complex elements per transform of the output (r2c: outer x (n/2 + 1);
        c2r: outer x n/2 pairs)"""
        return self.rows * (self.n // 2 + 1 if self.kind == "r2c" else self.n // 2)

    def exec_io_bytes(self):
        h, c = self.rows * (self.n // 2), self.rows * (self.n // 2 + 1)
        if self.kind == "r2c":
            return h * self.batch * ELEM_BYTES[self.in_fmt], c * self.batch * ELEM_BYTES[self.out_fmt]
        return c * self.batch * ELEM_BYTES[self.in_fmt], h * self.batch * ELEM_BYTES[self.out_fmt]


class LayoutPlan(Plan):
    """A plan reading and writing strided and/or planar layouts (Layout, or
    None for contiguous interleaved) on each side."""

    def __init__(self, dims, batch: int, in_layout: Layout | None, out_layout: Layout | None,
                 in_fmt: str = "fp16", out_fmt: str | None = None, scaling: str = "auto"):
        dims = [int(d) for d in (dims if hasattr(dims, "__len__") else [dims])]
        out_fmt = out_fmt or in_fmt
        self.dims, self.batch, self.in_fmt, self.out_fmt, self.scaling = dims, int(batch), in_fmt, out_fmt, scaling
        self.in_layout, self.out_layout = in_layout, out_layout
        self._h = ctypes.c_void_p()
        arr = (ctypes.c_longlong * len(dims))(*dims)
        ref = lambda l: ctypes.byref(l) if l is not None else None
        _check(lib().mcfft_plan_many_layout(ctypes.byref(self._h), len(dims), arr, self.batch, ref(in_layout),
                                            ref(out_layout), FORMATS[in_fmt], FORMATS[out_fmt], SCALINGS[scaling]),
               "mcfft_plan_many_layout")
        self.exponent = _static_exponent(self._h, scaling)
        self.n = int(np.prod(dims))

    def _side_bytes(self, lay, fmt):
        """Bytes of one array of a side: its extent in elements (interleaved)
        or scalars (planar: half an element)."""
        if lay is None:
            return self.n * self.batch * ELEM_BYTES[fmt]
        ext = (self.batch - 1) * lay.dist + sum((d - 1) * lay.strides[i] for i, d in enumerate(self.dims)) + 1
        return ext * ELEM_BYTES[fmt] // (2 if lay.planar else 1)

    def exec_io_bytes(self):
        return self._side_bytes(self.in_layout, self.in_fmt), self._side_bytes(self.out_layout, self.out_fmt)

    def exec_layout(self, in_re, in_im, out_re, out_im, direction: int = FORWARD, stream=None):
        """torch device tensors; *_im None for interleaved sides."""
        import torch
        ib, ob = self.exec_io_bytes()
        _check_tensor(in_re, ib, "input (re)")
        _check_tensor(out_re, ob, "output (re)")
        if in_im is not None:
            _check_tensor(in_im, ib, "input (im)")
        if out_im is not None:
            _check_tensor(out_im, ob, "output (im)")
        if stream is None:
            stream = torch.cuda.current_stream(in_re.device).cuda_stream
        p = lambda t: t.data_ptr() if t is not None else None
        _check(lib().mcfft_exec_layout(self._h, p(in_re), p(in_im), p(out_re), p(out_im), direction, stream),
               "mcfft_exec_layout")


def real_to_pairs(x: np.ndarray) -> np.ndarray:
    """batch x n reals -> complex pairs x[2m] + i x[2m+1] (the real storage layout)."""
    x = np.asarray(x, dtype=np.float64)
    return (x[..., 0::2] + 1j * x[..., 1::2]).reshape(-1)


def pairs_to_real(z: np.ndarray, n: int) -> np.ndarray:
    z = np.asarray(z).reshape(-1, n // 2)
    r = np.empty((z.shape[0], n))
    r[:, 0::2], r[:, 1::2] = z.real, z.imag
    return r


def error_norms(y, fmt: str, exponent: int, ref, stream=None, noalloc: bool = False):
    """(rel-L2, max-abs) of device tensor `y` (storage format `fmt`, times
    2^exponent) against `ref`, a device float64 tensor of interleaved complex
    values -- reduced on the GPU in FP64.  noalloc: use the entry point that
    keeps its partial sums in library-owned device memory (no allocation)."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream(y.device).cuda_stream
    n = ref.numel() // 2
    _check_tensor(y, n * ELEM_BYTES[fmt], "y")
    _check_tensor(ref, n * 16, "ref")
    r, m = ctypes.c_double(), ctypes.c_double()
    fn = lib().mcfft_error_norms_noalloc if noalloc else lib().mcfft_error_norms
    _check(fn(y.data_ptr(), FORMATS[fmt], exponent, ref.data_ptr(), n, ctypes.byref(r), ctypes.byref(m), stream),
           "mcfft_error_norms_noalloc" if noalloc else "mcfft_error_norms")
    return r.value, m.value


def components(fmt: str) -> int:
    return 4 if fmt == "split" else 2


def encode(x: np.ndarray, fmt: str) -> np.ndarray:
    """complex array -> host storage array (float16 / bfloat16-as-uint16 / split)."""
    x = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1)
    if fmt == "fp16":
        v = np.empty((x.size, 2), dtype=np.float16)
        v[:, 0], v[:, 1] = x.real, x.imag
        return v
    if fmt == "bf16":
        f = np.empty((x.size, 2), dtype=np.float32)
        f[:, 0], f[:, 1] = x.real, x.imag
        u = f.view(np.uint32)
        return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    hi = encode(x, "fp16")
    out = np.empty((x.size, 4), dtype=np.float16)
    out[:, 0:2] = hi
    out[:, 2] = (x.real - hi[:, 0].astype(np.float64)) * 2048.0
    out[:, 3] = (x.imag - hi[:, 1].astype(np.float64)) * 2048.0
    return out


def decode(v: np.ndarray, fmt: str, exponent: int = 0) -> np.ndarray:
    """host storage array -> complex128, times 2^exponent."""
    s = 2.0 ** exponent
    if fmt == "fp16":
        h = np.asarray(v).view(np.float16).reshape(-1, 2).astype(np.float64)
        return (h[:, 0] + 1j * h[:, 1]) * s
    if fmt == "bf16":
        u = np.asarray(v).view(np.uint16).reshape(-1, 2).astype(np.uint32) << 16
        f = u.view(np.float32).astype(np.float64)
        return (f[:, 0] + 1j * f[:, 1]) * s
    q = np.asarray(v).view(np.float16).reshape(-1, 4).astype(np.float64)
    return ((q[:, 0] + q[:, 2] / 2048.0) + 1j * (q[:, 1] + q[:, 3] / 2048.0)) * s


def to_device(x: np.ndarray, fmt: str, device="cuda"):
    """complex array -> torch device tensor in storage format."""
    import torch
    h = encode(x, fmt)
    if fmt == "bf16":
        return torch.from_numpy(h.view(np.int16)).to(device).view(torch.bfloat16)
    return torch.from_numpy(h).to(device)


def from_device(t, fmt: str, exponent: int = 0) -> np.ndarray:
    import torch
    if fmt == "bf16":
        h = t.view(torch.int16).cpu().numpy().view(np.uint16)
    else:
        h = t.cpu().numpy()
    return decode(h, fmt, exponent)


def empty(n_elems: int, fmt: str, device="cuda"):
    import torch
    dt = torch.bfloat16 if fmt == "bf16" else torch.float16
    return torch.empty((n_elems, components(fmt)), dtype=dt, device=device)
