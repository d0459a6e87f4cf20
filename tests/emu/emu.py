"""ctypes binding of the CPU emulation build (tests/emu/_build/libmcfft_emu.so)."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "libmcfft_emu.so")
_lib = None

FMT_F16, FMT_BF16, FMT_SPLIT = 0, 1, 2


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", _HERE, "-j8"], check=True)
        L = ctypes.CDLL(_LIB)
        L.emu_exec.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), ctypes.c_longlong, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_longlong, ctypes.POINTER(ctypes.c_int)]
        L.emu_exec.restype = ctypes.c_int
        L.emu_plan_describe.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), ctypes.c_longlong,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.emu_plan_describe.restype = ctypes.c_int
        _lib = L
    return _lib


def to_storage(x: np.ndarray, fmt: int) -> np.ndarray:
    """complex128 -> storage array (uint16 pairs, or hi/lo pairs for split)."""
    x = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1)
    if fmt == FMT_F16:
        v = np.empty(2 * x.size, dtype=np.float16)
        v[0::2] = x.real
        v[1::2] = x.imag
        return v.view(np.uint16)
    if fmt == FMT_BF16:
        f = np.empty(2 * x.size, dtype=np.float32)
        f[0::2] = x.real
        f[1::2] = x.imag
        u = f.view(np.uint32)
        r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
        return r
    hi = to_storage(x, FMT_F16).view(np.float16)
    re = x.real - hi[0::2].astype(np.float64)
    im = x.imag - hi[1::2].astype(np.float64)
    lo = np.empty(2 * x.size, dtype=np.float16)
    lo[0::2] = re * 2048.0
    lo[1::2] = im * 2048.0
    out = np.empty((x.size, 4), dtype=np.uint16)
    out[:, 0:2] = hi.view(np.uint16).reshape(-1, 2)
    out[:, 2:4] = lo.view(np.uint16).reshape(-1, 2)
    return out.reshape(-1)


def from_storage(v: np.ndarray, fmt: int) -> np.ndarray:
    v = np.asarray(v).view(np.uint16)
    if fmt == FMT_F16:
        h = v.view(np.float16).astype(np.float64)
        return h[0::2] + 1j * h[1::2]
    if fmt == FMT_BF16:
        f = (v.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        return f[0::2] + 1j * f[1::2]
    q = v.reshape(-1, 4).view(np.float16).astype(np.float64)
    return (q[:, 0] + q[:, 2] / 2048.0) + 1j * (q[:, 1] + q[:, 3] / 2048.0)


SCALE_NONE, SCALE_BALANCED, SCALE_FULL, SCALE_AUTO = 0, 1, 2, 3


def transform_exponents(batch: int) -> np.ndarray:
    """Per-transform output exponents the last emulated exec wrote (the
    runtime's mcfft_get_transform_exponents)."""
    L = lib()
    L.emu_transform_exponents.argtypes = [ctypes.c_void_p, ctypes.c_longlong]
    e = np.zeros(batch, dtype=np.int32)
    if L.emu_transform_exponents(e.ctypes.data, batch) != 0:
        raise RuntimeError("no exponents for this batch")
    return e


def status() -> int:
    """Status word of the last emulated exec (MCFFT_STATUS_* bits)."""
    L = lib()
    L.emu_status.restype = ctypes.c_uint
    return int(L.emu_status())


_input_exponents = None


def set_input_exponents(e):
    """Per-transform input exponents for the following execs (None: none)."""
    global _input_exponents
    L = lib()
    L.emu_set_input_exponents.argtypes = [ctypes.c_void_p, ctypes.c_longlong]
    if e is None:
        _input_exponents = None
        L.emu_set_input_exponents(None, 0)
        return
    a = np.ascontiguousarray(e, dtype=np.int32)
    _input_exponents = a.copy()
    L.emu_set_input_exponents(a.ctypes.data, a.size)


def _check_static(te, e, scaling):
    """static scaling: every transform's exponent is the plan's (plus its input exponent)"""
    if scaling != SCALE_AUTO:
        want = e + (_input_exponents if _input_exponents is not None else 0)
        assert (te == want).all(), (te, e)


def _scaled(y, batch):
    """stored values (batch rows) times 2^e of their transform"""
    e = transform_exponents(batch)
    return (np.asarray(y).reshape(batch, -1) * np.exp2(e.astype(np.float64))[:, None]).reshape(-1), e


def exec_fft(x, dims, batch, in_fmt, out_fmt, scaling=1, inverse=False, ws_limit=0):
    """Run the emulated plan; returns (complex128 output in true units, the
    per-transform output exponents).  Every plan writes one exponent per
    transform (the same for all of them under static scaling)."""
    src = to_storage(x, in_fmt)
    n = int(np.prod(dims)) * batch
    out = np.zeros(n * (4 if out_fmt == FMT_SPLIT else 2), dtype=np.uint16)
    d = (ctypes.c_longlong * len(dims))(*dims)
    e = ctypes.c_int(0)
    rc = lib().emu_exec(len(dims), d, batch, in_fmt, out_fmt, scaling, int(inverse), src.ctypes.data, out.ctypes.data,
                        ws_limit, ctypes.byref(e))
    if rc != 0:
        raise RuntimeError(f"emu_exec status {rc}")
    y, te = _scaled(from_storage(out, out_fmt), batch)
    _check_static(te, e.value, scaling)
    return y, te


def exec_ptr(in_ptr, out_ptr, dims, batch, in_fmt, out_fmt, scaling=1, inverse=False, only_pass=None,
             blocks=None):
    """Run the emulated plan on caller-owned host memory (raw addresses of
    buffers as large as the plan's input and output).  only_pass / blocks
    (first, count): run only that pass, launched with one workgroup per tile,
    and of it only those workgroups.  Returns the plan's output exponent."""
    L = lib()
    L.emu_restrict.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong]
    d = (ctypes.c_longlong * len(dims))(*dims)
    e = ctypes.c_int(0)
    first, count = blocks if blocks is not None else (0, 1 << 62)
    L.emu_restrict(-1 if only_pass is None else int(only_pass), first, count)
    try:
        rc = L.emu_exec(len(dims), d, batch, in_fmt, out_fmt, scaling, int(inverse), in_ptr, out_ptr, 0,
                        ctypes.byref(e))
    finally:
        L.emu_restrict(-1, 0, 0)
    if rc != 0:
        raise RuntimeError(f"emu_exec status {rc}")
    return e.value


def describe(dims, batch, in_fmt, out_fmt, scaling=1):
    buf = ctypes.create_string_buffer(8192)
    d = (ctypes.c_longlong * len(dims))(*dims)
    rc = lib().emu_plan_describe(len(dims), d, batch, in_fmt, out_fmt, scaling, buf, 8192)
    if rc != 0:
        raise RuntimeError(f"plan status {rc}")
    return buf.value.decode()


def structures(dims, batch, in_fmt, out_fmt, scaling=1) -> list:
    """The plan structures the autotuner tries for this shape (opts strings,
    "R=<outer radix>;L2=<four-step log2 N2>"), the default first."""
    L = lib()
    L.emu_plan_structures.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), ctypes.c_longlong,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    buf = ctypes.create_string_buffer(4096)
    d = (ctypes.c_longlong * len(dims))(*dims)
    L.emu_plan_structures(len(dims), d, batch, in_fmt, out_fmt, scaling, buf, 4096)
    return buf.value.decode().split()


def describe_opts(dims, batch, in_fmt, out_fmt, scaling=1, outer_radix=-1, fourstep_lg2=-1) -> str:
    L = lib()
    L.emu_plan_describe_opts.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), ctypes.c_longlong,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_char_p, ctypes.c_int]
    buf = ctypes.create_string_buffer(8192)
    d = (ctypes.c_longlong * len(dims))(*dims)
    rc = L.emu_plan_describe_opts(len(dims), d, batch, in_fmt, out_fmt, scaling, outer_radix, fourstep_lg2, buf, 8192)
    if rc != 0:
        raise RuntimeError(f"plan status {rc}")
    return buf.value.decode()


def set_structure(outer_radix=-1, fourstep_lg2=-1):
    """Structure of the plans of later exec_fft / exec_ptr calls (-1: defaults)."""
    L = lib()
    L.emu_set_structure.argtypes = [ctypes.c_int, ctypes.c_int]
    L.emu_set_structure(outer_radix, fourstep_lg2)


def tune_cache_lookup(text: str, build_id: str, key: str):
    """The tuning-cache entry for `key` as a library built as `build_id` reads
    the cache file `text` (re-formatted), or None."""
    L = lib()
    L.emu_tune_cache_lookup.argtypes = [ctypes.c_char_p] * 3 + [ctypes.c_char_p, ctypes.c_int]
    buf = ctypes.create_string_buffer(4096)
    rc = L.emu_tune_cache_lookup(text.encode(), build_id.encode(), key.encode(), buf, 4096)
    return buf.value.decode() if rc == 0 else None


def exec_gathered(buf_in, buf_out, dims, in_offs, out_offs, in_fmt, out_fmt, scaling=1, inverse=False, ws_limit=0):
    """A gathered batched plan on storage arrays: transform t reads its
    prod(dims) elements at element in_offs[t] of buf_in and writes them at
    out_offs[t] of buf_out (the layout a ragged plan's same-shape items have).
    Returns the per-transform output exponents."""
    L = lib()
    L.emu_set_gather.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_longlong,
                                 ctypes.c_longlong]
    io = np.ascontiguousarray(in_offs, dtype=np.int64)
    oo = np.ascontiguousarray(out_offs, dtype=np.int64)
    comp = lambda fmt: 4 if fmt == FMT_SPLIT else 2
    L.emu_set_gather(io.ctypes.data, oo.ctypes.data, len(io), buf_in.size // comp(in_fmt),
                     buf_out.size // comp(out_fmt))
    d = (ctypes.c_longlong * len(dims))(*dims)
    e = ctypes.c_int(0)
    try:
        rc = L.emu_exec(len(dims), d, len(io), in_fmt, out_fmt, scaling, int(inverse), buf_in.ctypes.data,
                        buf_out.ctypes.data, ws_limit, ctypes.byref(e))
    finally:
        L.emu_set_gather(None, None, 0, 0, 0)
    if rc != 0:
        raise RuntimeError(f"emu_exec status {rc}")
    return transform_exponents(len(io))


def ragged_exec(xs, shapes, in_fmt, out_fmt, layout=None, scaling=1, inverse=False, nblocks=3, single=False):
    """Ragged batch: xs = list of complex arrays, shapes = list of dims lists.
    layout: list of (in_off, out_off) element offsets (default back to back).
    Returns (list of outputs * 2^exponent, status)."""
    n_items = len(xs)
    sizes = [int(np.prod(s)) for s in shapes]
    if layout is None:
        offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        layout = [(int(o), int(o)) for o in offs]
    in_total = max(o[0] + s for o, s in zip(layout, sizes))
    out_total = max(o[1] + s for o, s in zip(layout, sizes))
    src_c = np.zeros(in_total, dtype=np.complex128)
    for x, (io, _), s in zip(xs, layout, sizes):
        src_c[io:io + s] = x
    src = to_storage(src_c, in_fmt)
    out = np.zeros(out_total * (4 if out_fmt == FMT_SPLIT else 2), dtype=np.uint16)
    spec = np.zeros((n_items, 6), dtype=np.int64)
    for i, (sh, (io, oo)) in enumerate(zip(shapes, layout)):
        spec[i, 0] = len(sh)
        spec[i, 1:1 + len(sh)] = sh
        for d in range(len(sh), 3):
            spec[i, 1 + d] = 1
        spec[i, 4], spec[i, 5] = io, oo
    exps = np.zeros(n_items, dtype=np.int32)
    st = ctypes.c_uint(0)
    L = lib()
    L.emu_ragged_exec.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint),
                                  ctypes.c_int]
    L.emu_ragged_exec.restype = ctypes.c_int
    rc = L.emu_ragged_exec(n_items, spec.ctypes.data, in_fmt, out_fmt, scaling, int(inverse), src.ctypes.data,
                           out.ctypes.data, nblocks, exps.ctypes.data, ctypes.byref(st), int(single))
    if rc != 0:
        raise RuntimeError(f"emu_ragged_exec status {rc}")
    y = from_storage(out, out_fmt)
    outs = [y[oo:oo + s] * (2.0 ** e) for (io, oo), s, e in zip(layout, sizes, exps)]
    return outs, st.value


def ragged_check(shapes, in_fmt, out_fmt, scaling=1):
    """Structural checks of the ragged plan of `shapes` laid out back to back
    (no kernel runs).  Returns (violations, stats)."""
    sizes = [int(np.prod(s)) for s in shapes]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    spec = np.zeros((len(shapes), 6), dtype=np.int64)
    for i, s in enumerate(shapes):
        spec[i, 0] = len(s)
        spec[i, 1:1 + len(s)] = s
        for d in range(len(s), 3):
            spec[i, 1 + d] = 1
        spec[i, 4] = spec[i, 5] = offs[i]
    st = np.zeros(6, dtype=np.int64)
    L = lib()
    L.emu_ragged_check.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_void_p]
    L.emu_ragged_check.restype = ctypes.c_longlong
    bad = L.emu_ragged_check(len(shapes), spec.ctypes.data, in_fmt, out_fmt, scaling, st.ctypes.data)
    if bad < 0:
        raise RuntimeError(f"ragged_build status {-bad}")
    keys = ("tasks", "passes", "phases", "counters", "ws_elems", "tw_entries")
    return int(bad), {k: int(v) for k, v in zip(keys, st)}


def set_kernel_override(name: str | None):
    """Run `name` for every pass it can serve (None: the planner's choice)."""
    L = lib()
    L.emu_set_kernel_override.argtypes = [ctypes.c_char_p]
    L.emu_set_kernel_override((name or "").encode())


def last_kernels() -> str:
    """Kernel names of the passes of the last exec_fft call."""
    L = lib()
    L.emu_last_kernels.restype = ctypes.c_char_p
    return L.emu_last_kernels().decode()


def real_exec(kind: str, data, n: int, batch: int, in_fmt: int, out_fmt: int, scaling: int = 1):
    """kind "r2c": data = batch x n reals -> batch x (n/2+1) complex (times 2^e);
    kind "c2r": data = batch x (n/2+1) complex -> batch x n reals (times 2^e)."""
    h = n // 2
    if kind == "r2c":
        x = np.asarray(data, dtype=np.float64).reshape(batch, n)
        src = to_storage((x[:, 0::2] + 1j * x[:, 1::2]).reshape(-1), in_fmt)
        out = np.zeros((h + 1) * batch * (4 if out_fmt == FMT_SPLIT else 2), dtype=np.uint16)
    else:
        src = to_storage(np.asarray(data, dtype=np.complex128).reshape(-1), in_fmt)
        out = np.zeros(h * batch * (4 if out_fmt == FMT_SPLIT else 2), dtype=np.uint16)
    e = ctypes.c_int(0)
    L = lib()
    L.emu_real_exec.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    L.emu_real_exec.restype = ctypes.c_int
    rc = L.emu_real_exec(1 if kind == "r2c" else 2, n, batch, in_fmt, out_fmt, scaling, src.ctypes.data,
                         out.ctypes.data, ctypes.byref(e))
    if rc != 0:
        raise RuntimeError(f"emu_real_exec status {rc}")
    y, te = _scaled(from_storage(out, out_fmt), batch)
    _check_static(te, e.value, scaling)
    if kind == "r2c":
        return y.reshape(batch, h + 1)
    z = y.reshape(batch, h)
    r = np.empty((batch, n))
    r[:, 0::2], r[:, 1::2] = z.real, z.imag
    return r


def real_exec_nd(kind: str, data, dims, batch: int, in_fmt: int, out_fmt: int, scaling: int = 1):
    """Real transforms of shape `dims` (last dimension n): kind "r2c": data =
    batch x dims reals -> batch x outer x (n/2+1) complex; kind "c2r": data =
    batch x outer x (n/2+1) complex -> batch x dims reals.  True units (output
    exponents applied)."""
    dims = list(dims)
    n, h = dims[-1], dims[-1] // 2
    outer = int(np.prod(dims[:-1])) if len(dims) > 1 else 1
    if kind == "r2c":
        x = np.asarray(data, dtype=np.float64).reshape(batch * outer, n)
        src = to_storage((x[:, 0::2] + 1j * x[:, 1::2]).reshape(-1), in_fmt)
        out = np.zeros(outer * (h + 1) * batch * (4 if out_fmt == FMT_SPLIT else 2), dtype=np.uint16)
    else:
        src = to_storage(np.asarray(data, dtype=np.complex128).reshape(-1), in_fmt)
        out = np.zeros(outer * h * batch * (4 if out_fmt == FMT_SPLIT else 2), dtype=np.uint16)
    e = ctypes.c_int(0)
    L = lib()
    L.emu_real_exec_nd.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), ctypes.c_longlong,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_int)]
    d = (ctypes.c_longlong * len(dims))(*dims)
    rc = L.emu_real_exec_nd(1 if kind == "r2c" else 2, len(dims), d, batch, in_fmt, out_fmt, scaling,
                            src.ctypes.data, out.ctypes.data, ctypes.byref(e))
    if rc != 0:
        raise RuntimeError(f"emu_real_exec_nd status {rc}")
    y, te = _scaled(from_storage(out, out_fmt), batch)
    _check_static(te, e.value, scaling)
    if kind == "r2c":
        return y.reshape([batch] + dims[:-1] + [h + 1])
    z = y.reshape(batch * outer, h)
    r = np.empty((batch * outer, n))
    r[:, 0::2], r[:, 1::2] = z.real, z.imag
    return r.reshape([batch] + dims)


def layout_encode(x, fmt: int, planar: int, positions, extent: int, fill: int = 0x7E00):
    """Place complex values x[k] at element positions[k] of a caller-side buffer
    of `extent` elements: interleaved -> one uint16 array (2 or 4 components per
    element); planar -> (re, im) arrays of scalars (1 or 2 uint16 per scalar).
    Unused positions hold `fill` (a NaN pattern) so stray reads show."""
    v = to_storage(np.asarray(x).reshape(-1), fmt).reshape(-1, 4 if fmt == FMT_SPLIT else 2)
    pos = np.asarray(positions, dtype=np.int64).reshape(-1)
    if not planar:
        buf = np.full((extent, v.shape[1]), fill, dtype=np.uint16)
        buf[pos] = v
        return buf.reshape(-1), None
    w = 2 if fmt == FMT_SPLIT else 1
    re = np.full((extent, w), fill, dtype=np.uint16)
    im = np.full((extent, w), fill, dtype=np.uint16)
    if fmt == FMT_SPLIT:
        re[pos] = v[:, [0, 2]]
        im[pos] = v[:, [1, 3]]
    else:
        re[pos, 0] = v[:, 0]
        im[pos, 0] = v[:, 1]
    return re.reshape(-1), im.reshape(-1)


def layout_decode(re, im, fmt: int, planar: int, positions) -> np.ndarray:
    """Inverse of layout_encode at `positions` (complex128, unscaled)."""
    pos = np.asarray(positions, dtype=np.int64).reshape(-1)
    c = 4 if fmt == FMT_SPLIT else 2
    if not planar:
        return from_storage(re.reshape(-1, c)[pos].reshape(-1), fmt)
    w = c // 2
    r, i = re.reshape(-1, w)[pos], im.reshape(-1, w)[pos]
    v = np.empty((pos.size, c), dtype=np.uint16)
    if fmt == FMT_SPLIT:
        v[:, [0, 2]] = r
        v[:, [1, 3]] = i
    else:
        v[:, 0], v[:, 1] = r[:, 0], i[:, 0]
    return from_storage(v.reshape(-1), fmt)


def layout_positions(dims, batch, strides, dist) -> np.ndarray:
    """Element offset of every (b, i_0, ..) in row-major order."""
    idx = np.zeros((batch,) + tuple(dims), dtype=np.int64)
    idx += (np.arange(batch, dtype=np.int64) * dist).reshape((batch,) + (1,) * len(dims))
    for d, (nd, sd) in enumerate(zip(dims, strides)):
        shape = [1] * (len(dims) + 1)
        shape[d + 1] = nd
        idx += (np.arange(nd, dtype=np.int64) * sd).reshape(shape)
    return idx.reshape(-1)


def layout_set_scalar_only(v: bool):
    """Force the one-element-per-thread gather/scatter path (test hook)."""
    lib().emu_layout_set_scalar_only(1 if v else 0)


def layout_vec_used() -> bool:
    """Whether the last layout_exec used the four-element path on a side."""
    return bool(lib().emu_layout_vec_used())


def layout_exec(x, dims, batch, in_layout, out_layout, in_fmt, out_fmt, scaling=1, inverse=False):
    """x: batch x prod(dims) complex.  in_layout / out_layout: (strides, dist,
    planar).  Returns (y as batch x prod(dims) complex times 2^e, output buffers,
    output positions) so callers can also check untouched positions."""
    rank = len(dims)
    L = lib()
    LL3 = ctypes.c_longlong * 3
    vp = ctypes.c_void_p
    L.emu_layout_exec.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), ctypes.c_longlong,
                                  ctypes.POINTER(ctypes.c_longlong), ctypes.c_longlong, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_longlong), ctypes.c_longlong, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp,
                                  ctypes.POINTER(ctypes.c_int)]
    L.emu_layout_exec.restype = ctypes.c_int
    (is_, idist, ipl), (os_, odist, opl) = in_layout, out_layout
    ipos = layout_positions(dims, batch, is_, idist)
    opos = layout_positions(dims, batch, os_, odist)
    iext = int(ipos.max()) + 1
    oext = int(opos.max()) + 1
    in_re, in_im = layout_encode(x, in_fmt, ipl, ipos, iext)
    out_re, out_im = layout_encode(np.zeros(0), out_fmt, opl, np.zeros(0, dtype=np.int64), oext, fill=0x7C01)
    e = ctypes.c_int(0)
    ptr = lambda a: a.ctypes.data if a is not None else None
    rc = L.emu_layout_exec(rank, (ctypes.c_longlong * rank)(*dims), batch, LL3(*(list(is_) + [0] * (3 - rank))),
                           idist, ipl, LL3(*(list(os_) + [0] * (3 - rank))), odist, opl, in_fmt, out_fmt, scaling,
                           1 if inverse else 0, ptr(in_re), ptr(in_im), ptr(out_re), ptr(out_im), ctypes.byref(e))
    if rc != 0:
        raise RuntimeError(f"emu_layout_exec status {rc}")
    y, te = _scaled(layout_decode(out_re, out_im, out_fmt, opl, opos), batch)
    _check_static(te, e.value, scaling)
    return y.reshape(batch, -1), (out_re, out_im), opos
