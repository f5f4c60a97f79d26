"""ctypes binding of the CPU oracle (oracle/oracle.cpp) — test infrastructure.

The oracle is a single-threaded C++ restatement of ph-core's algorithms; it is
the checker for the device path and never part of it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

MAXR = 8
DT_I32, DT_F32, DT_F64, DT_BF16, DT_BOOL = 0, 1, 2, 3, 4
NP_OF = {DT_I32: np.int32, DT_F32: np.float32, DT_F64: np.float64, DT_BF16: np.uint16, DT_BOOL: np.uint8}
DT_OF_NP = {np.dtype(np.int32): DT_I32, np.dtype(np.float32): DT_F32, np.dtype(np.float64): DT_F64,
            np.dtype(np.uint16): DT_BF16, np.dtype(np.uint8): DT_BOOL, np.dtype(np.bool_): DT_BOOL}

E_INDEX, E_SHAPE, E_DIMENSION, E_ARGUMENT, E_OVERFLOW, E_DIVZERO, E_GENERIC = -1, -2, -3, -4, -5, -6, -7


class Lit(C.Structure):
    _fields_ = [("kind", C.c_int32), ("has_first", C.c_int32), ("has_last", C.c_int32),
                ("has_step", C.c_int32), ("first", C.c_int64), ("last", C.c_int64),
                ("step", C.c_int64), ("exclusive", C.c_int32), ("pad_", C.c_int32)]


class Region(C.Structure):
    _fields_ = [("rank", C.c_int32), ("drop", C.c_int32), ("first", C.c_int64 * MAXR),
                ("step", C.c_int64 * MAXR), ("last", C.c_int64 * MAXR),
                ("proper_shape", C.c_int64 * MAXR), ("degeneracy", C.c_int32 * MAXR),
                ("reduced_rank", C.c_int32), ("pad_", C.c_int32),
                ("reduced_shape", C.c_int64 * MAXR)]

    def get(self, name):
        n = self.reduced_rank if name == "reduced_shape" else self.rank
        return [int(v) for v in getattr(self, name)[:n]]

    @property
    def shape(self):
        return self.get("reduced_shape")

    @property
    def size(self):
        s = 1
        for v in self.shape:
            s *= v
        return s if self.reduced_rank else 0


class Transform(C.Structure):
    _fields_ = [("kind", C.c_int32), ("rank_a", C.c_int32), ("a", C.c_int64 * MAXR),
                ("rank_b", C.c_int32), ("pad_", C.c_int32), ("b", C.c_int64 * MAXR),
                ("region", Region)]


class Input(C.Structure):
    _fields_ = [("buf", C.c_void_p), ("dtype", C.c_int32), ("rank", C.c_int32), ("shape", C.c_int64 * MAXR),
                ("chain", C.POINTER(Transform)), ("nchain", C.c_int32), ("pad_", C.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        i64p = C.POINTER(C.c_int64)
        L.or_infer_range.argtypes = [C.POINTER(Lit), C.c_int64, i64p, C.c_char_p, C.c_int]
        L.or_canonicalize_range.argtypes = [C.POINTER(Lit), C.c_int64, i64p, C.c_char_p, C.c_int]
        L.or_region_new.argtypes = [C.POINTER(Lit), C.c_int, i64p, C.c_int, C.c_int, C.POINTER(Region), C.c_char_p, C.c_int]
        L.or_region_new_absolute.argtypes = [C.POINTER(Lit), C.c_int, C.c_int, C.POINTER(Region), C.c_char_p, C.c_int]
        L.or_region_new_trim.argtypes = [C.POINTER(Lit), C.c_int, i64p, C.c_int, i64p, C.c_int, C.POINTER(Region), C.c_char_p, C.c_int]
        L.or_region_cover.argtypes = [i64p, C.c_int, C.c_int, C.POINTER(Region)]
        L.or_region_trim.argtypes = [C.POINTER(Region), i64p, C.c_int, C.c_char_p, C.c_int]
        L.or_region_reverse.argtypes = [C.POINTER(Region)]
        L.or_region_translate.argtypes = [C.POINTER(Region), i64p, C.c_int, C.c_char_p, C.c_int]
        L.or_region_includes.argtypes = [C.POINTER(Region), i64p, C.c_int]
        L.or_region_fits_in.argtypes = [C.POINTER(Region), i64p, C.c_int, C.c_char_p, C.c_int]
        L.or_region_local_to_absolute.argtypes = [C.POINTER(Region), i64p, i64p]
        L.or_gather.argtypes = [C.c_void_p, i64p, C.c_int, C.c_int, C.POINTER(Region), C.c_void_p]
        L.or_gather.restype = C.c_int64
        L.or_scatter.argtypes = [C.c_void_p, i64p, C.c_int, C.c_int, C.POINTER(Region), C.c_void_p, C.c_int64]
        L.or_scatter.restype = C.c_int64
        L.or_fill_region.argtypes = [C.c_void_p, i64p, C.c_int, C.c_int, C.POINTER(Region), C.c_void_p]
        L.or_fill_region.restype = C.c_int64
        L.or_chain_apply.argtypes = [C.POINTER(Transform), C.c_int, i64p, C.c_int, i64p]
        L.or_view_gather.argtypes = [C.c_void_p, i64p, C.c_int, C.c_int, C.POINTER(Transform), C.c_int, i64p, C.c_int, C.c_void_p]
        L.or_view_gather.restype = C.c_int64
        L.or_view_scatter.argtypes = [C.c_void_p, i64p, C.c_int, C.c_int, C.POINTER(Transform), C.c_int, C.POINTER(Region), C.c_void_p, C.c_int64]
        L.or_view_scatter.restype = C.c_int64
        L.or_tile.argtypes = [C.c_void_p, i64p, C.c_int, C.c_int, i64p, C.c_void_p]
        L.or_set_mask.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        L.or_set_mask.restype = None
        L.or_tile.restype = C.c_int64
        L.or_binop.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int64, C.c_void_p, C.POINTER(C.c_int)]
        L.or_unop.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(C.c_int)]
        L.or_eval.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_void_p), C.c_int64, C.c_void_p, C.POINTER(C.c_int)]
        L.or_set_chunk_lazy.argtypes = [C.c_void_p, i64p, C.c_int, C.POINTER(Transform), C.c_int,
                                        C.POINTER(Region), C.c_char_p, C.c_int, C.POINTER(Input), i64p, C.c_int,
                                        C.POINTER(C.c_int)]
        L.or_set_chunk_lazy.restype = C.c_int64
        L.or_set_mask_lazy.argtypes = [C.c_void_p, C.c_int, i64p, C.c_int, C.c_char_p, C.c_char_p, C.c_void_p,
                                       C.c_int, C.POINTER(Input), C.POINTER(C.c_int)]
        L.or_sum_seq_f64_from.argtypes = [C.c_void_p, C.c_int64, C.c_double]
        L.or_sum_seq_f64_from.restype = C.c_double
        L.or_sum_seq_f64.argtypes = [C.c_void_p, C.c_int64]
        L.or_sum_seq_f64.restype = C.c_double
        L.or_concatenate.argtypes = [C.c_int, C.POINTER(C.c_void_p), i64p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.or_concatenate.restype = C.c_int64
        L.or_sum_seq_f32.argtypes = [C.c_void_p, C.c_int64]
        L.or_sum_seq_f32.restype = C.c_float
        L.or_batch_loop.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        L.or_max_seq_f64.argtypes = [C.c_void_p, C.c_int64]
        L.or_max_seq_f64.restype = C.c_double
        L.or_sum_exact.argtypes = [C.c_void_p, C.c_int64]
        L.or_sum_exact.restype = C.c_double
        L.or_sum_i32.argtypes = [C.c_void_p, C.c_int64, i64p]
        L.or_slices_reduce.argtypes = [C.c_void_p, i64p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.or_heat1d_reference.argtypes = [C.c_int64, C.c_void_p]
        L.or_heat2d.argtypes = [C.c_int, C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_int64, C.c_int64, C.c_void_p]
        L.or_fill_random.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.c_uint64, C.c_int, C.c_double, C.c_double]
        L.or_draw.argtypes = [C.c_uint64, C.c_uint64]
        L.or_draw.restype = C.c_uint64
        _lib = L
    return _lib


class OracleError(Exception):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _i64(seq):
    arr = (C.c_int64 * max(1, len(seq)))(*[int(v) for v in seq])
    return arr


def _check(rc, buf):
    if rc != 0:
        raise OracleError(rc, buf.value.decode(errors="replace"))


def make_lit(parsed):
    """parsed: ('index', i) or ('range', first, last, step, exclusive) with None for nil."""
    L = Lit()
    if parsed[0] == "index":
        L.kind = 1
        L.first = parsed[1]
        return L
    _, f, l, s, e = parsed
    L.kind = 0
    L.has_first, L.first = (f is not None), (f or 0)
    L.has_last, L.last = (l is not None), (l or 0)
    L.has_step, L.step = (s is not None), (s or 0)
    L.exclusive = 1 if e else 0
    return L


def lits(parsed_list):
    arr = (Lit * max(1, len(parsed_list)))()
    for i, p in enumerate(parsed_list):
        arr[i] = make_lit(p)
    return arr


def infer_range(parsed, bound):
    out = (C.c_int64 * 4)()
    buf = C.create_string_buffer(256)
    _check(lib().or_infer_range(C.byref(make_lit(parsed)), bound, out, buf, 256), buf)
    return tuple(int(v) for v in out)


def canonicalize_range(parsed, bound):
    out = (C.c_int64 * 4)()
    buf = C.create_string_buffer(256)
    _check(lib().or_canonicalize_range(C.byref(make_lit(parsed)), bound, out, buf, 256), buf)
    return tuple(int(v) for v in out)


def region(parsed_list, bound, drop=True):
    r = Region()
    buf = C.create_string_buffer(256)
    _check(lib().or_region_new(lits(parsed_list), len(parsed_list), _i64(bound), len(bound), int(drop), C.byref(r), buf, 256), buf)
    return r


def region_absolute(parsed_list, drop=True):
    r = Region()
    buf = C.create_string_buffer(256)
    _check(lib().or_region_new_absolute(lits(parsed_list), len(parsed_list), int(drop), C.byref(r), buf, 256), buf)
    return r


def region_trimmed(parsed_list, bound, drop, trim_to):
    r = Region()
    buf = C.create_string_buffer(256)
    b = _i64(bound) if bound is not None else None
    _check(lib().or_region_new_trim(lits(parsed_list), len(parsed_list), b, int(drop), _i64(trim_to), len(trim_to), C.byref(r), buf, 256), buf)
    return r


def region_cover(shape, drop=True):
    r = Region()
    lib().or_region_cover(_i64(shape), len(shape), int(drop), C.byref(r))
    return r


def region_trim(r, bound):
    buf = C.create_string_buffer(256)
    _check(lib().or_region_trim(C.byref(r), _i64(bound), len(bound), buf, 256), buf)
    return r


def region_reverse(r):
    lib().or_region_reverse(C.byref(r))
    return r


def region_translate(r, offset):
    buf = C.create_string_buffer(256)
    _check(lib().or_region_translate(C.byref(r), _i64(offset), len(offset), buf, 256), buf)
    return r


def region_includes(r, coord):
    return bool(lib().or_region_includes(C.byref(r), _i64(coord), len(coord)))


def region_fits_in(r, bound):
    buf = C.create_string_buffer(256)
    rc = lib().or_region_fits_in(C.byref(r), _i64(bound), len(bound), buf, 256)
    if rc < 0:
        raise OracleError(rc, buf.value.decode())
    return bool(rc)


def local_to_absolute(r, local):
    out = (C.c_int64 * MAXR)()
    lib().or_region_local_to_absolute(C.byref(r), _i64(local), out)
    return [int(v) for v in out[:r.rank]]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def gather(src, r):
    """NArray#unsafe_fetch_chunk: src is a numpy array (its shape is the NArray's)."""
    src = np.ascontiguousarray(src)
    out = np.zeros(max(r.size, 1), dtype=src.dtype)
    n = lib().or_gather(_ptr(src), _i64(src.shape), src.ndim, src.itemsize, C.byref(r), _ptr(out))
    return out[:n].reshape(r.shape) if r.size else out[:0]


def scatter(dst, r, src):
    """NArray#unsafe_set_chunk(region, src) in place on numpy dst."""
    src = np.ascontiguousarray(src, dtype=dst.dtype).reshape(-1)
    lib().or_scatter(_ptr(dst), _i64(dst.shape), dst.ndim, dst.itemsize, C.byref(r), _ptr(src), src.size)
    return dst


def fill_region(dst, r, value):
    v = np.array([value], dtype=dst.dtype)
    lib().or_fill_region(_ptr(dst), _i64(dst.shape), dst.ndim, dst.itemsize, C.byref(r), _ptr(v))
    return dst


# transform chain: list of ("region", Region) | ("reshape", src_shape, new_shape)
# | ("permute", pattern) | ("reverse", shape), newest FIRST
def _chain(chain):
    arr = (Transform * max(1, len(chain)))()
    for i, t in enumerate(chain):
        T = arr[i]
        kind = t[0]
        if kind == "region":
            T.kind = 0
            T.region = t[1]
        elif kind == "reshape":
            T.kind = 1
            T.rank_a = len(t[1])
            T.rank_b = len(t[2])
            for k, v in enumerate(t[1]):
                T.a[k] = v
            for k, v in enumerate(t[2]):
                T.b[k] = v
        elif kind == "permute":
            T.kind = 2
            T.rank_a = len(t[1])
            for k, v in enumerate(t[1]):
                T.a[k] = v
        elif kind == "reverse":
            T.kind = 3
            T.rank_a = len(t[1])
            for k, v in enumerate(t[1]):
                T.a[k] = v
    return arr


def chain_apply(chain, coord):
    out = (C.c_int64 * MAXR)()
    n = lib().or_chain_apply(_chain(chain), len(chain), _i64(coord), len(coord), out)
    return [int(v) for v in out[:n]]


def view_gather(src, chain, view_shape):
    src = np.ascontiguousarray(src)
    n = int(np.prod(view_shape)) if view_shape else 0
    out = np.zeros(max(n, 1), dtype=src.dtype)
    lib().or_view_gather(_ptr(src), _i64(src.shape), src.ndim, src.itemsize, _chain(chain), len(chain),
                         _i64(view_shape), len(view_shape), _ptr(out))
    return out[:n].reshape(view_shape)


def view_scatter(dst, chain, r, src):
    src = np.ascontiguousarray(src, dtype=dst.dtype).reshape(-1)
    lib().or_view_scatter(_ptr(dst), _i64(dst.shape), dst.ndim, dst.itemsize, _chain(chain), len(chain),
                          C.byref(r), _ptr(src), src.size)
    return dst


def _inputs(inputs, keep):
    """Lazily read operands: each is (array, chain newest first); the array
    is read where it lies, so passing the destination array itself makes the
    expression see the writes made so far.  `keep` holds the ctypes buffers
    alive for the call."""
    arr = (Input * max(1, len(inputs)))()
    for k, (a, chain) in enumerate(inputs):
        assert a.flags.c_contiguous
        ch = _chain(chain)
        keep.append(ch)
        I = arr[k]
        I.buf = a.ctypes.data
        I.dtype = _dt_of(a)
        I.rank = a.ndim
        for i, v in enumerate(a.shape):
            I.shape[i] = v
        I.chain = C.cast(ch, C.POINTER(Transform))
        I.nchain = len(chain)
    return arr


def set_chunk_lazy(dst, dchain, r, expr, inputs, src_shape):
    """MultiWritable#set_chunk with a lazily evaluated source (a View is the
    expression "x0", a ProcView its block): in place on `dst`; returns the
    flags raised."""
    assert dst.flags.c_contiguous
    keep = []
    flags = C.c_int(0)
    n = lib().or_set_chunk_lazy(_ptr(dst), _i64(dst.shape), dst.ndim, _chain(dchain), len(dchain), C.byref(r),
                                expr.encode(), len(inputs), _inputs(inputs, keep), _i64(src_shape), len(src_shape),
                                C.byref(flags))
    if n < 0:
        raise ValueError("oracle could not parse " + expr)
    return flags.value


def set_mask_lazy(dst, mask_expr, value, inputs):
    """NArray#[]=(mask, value) with the mask and the value (an expression
    string, or a scalar) evaluated lazily in the reference's loop order; in
    place on `dst`; returns the flags raised."""
    assert dst.flags.c_contiguous
    keep = []
    flags = C.c_int(0)
    if isinstance(value, str):
        vexpr, sc = value.encode(), None
    else:
        vexpr, sc = None, np.array([value], dtype=dst.dtype)
        keep.append(sc)
    rc = lib().or_set_mask_lazy(_ptr(dst), _dt_of(dst), _i64(dst.shape), dst.ndim, mask_expr.encode(), vexpr,
                                None if sc is None else sc.ctypes.data, len(inputs), _inputs(inputs, keep),
                                C.byref(flags))
    if rc < 0:
        raise ValueError("oracle could not parse an expression")
    return flags.value


def set_mask(dst, mask, value):
    """NArray#[]=(mask, value) in place on a contiguous numpy array; `value`
    is a scalar or an array of dst's shape."""
    assert dst.flags.c_contiguous and mask.shape == dst.shape
    m = np.ascontiguousarray(mask, dtype=np.bool_).view(np.uint8)
    is_arr = isinstance(value, np.ndarray)
    v = np.ascontiguousarray(value, dtype=dst.dtype) if is_arr else np.array([value], dtype=dst.dtype)
    lib().or_set_mask(_ptr(dst), dst.size, dst.itemsize, _ptr(m), _ptr(v), int(is_arr))
    return dst


def tile(src, counts):
    src = np.ascontiguousarray(src)
    shape = [s * c for s, c in zip(src.shape, counts)]
    out = np.zeros(max(1, int(np.prod(shape))), dtype=src.dtype)
    n = lib().or_tile(_ptr(src), _i64(src.shape), src.ndim, src.itemsize, _i64(counts), _ptr(out))
    return out[:n].reshape(shape)


OPS = ["add", "sub", "mul", "div", "fdiv", "mod", "pow", "wadd", "wsub", "wmul", "wpow", "and", "or",
       "xor", "gt", "lt", "ge", "le", "eq", "ne", "cmp", "neg", "pos", "bnot", "sqrt", "abs"]


def _dt_of(a):
    return DT_OF_NP[np.dtype(a.dtype)]


def eval_expr(expr, inputs, bf16=()):
    """Evaluate a prefix expression elementwise; inputs are equal-size numpy
    arrays (uint16 arrays listed in `bf16` are BF16 storage).  Returns
    (result array, flags)."""
    n = inputs[0].size
    dts = []
    arrs = []
    for k, a in enumerate(inputs):
        a = np.ascontiguousarray(a).reshape(-1)
        if a.dtype == np.bool_:
            a = a.astype(np.uint8)
        dts.append(DT_BF16 if k in bf16 else _dt_of(a))
        arrs.append(a)
    out = np.zeros(max(n, 1) * 8, dtype=np.uint8)
    flags = C.c_int(0)
    dt_arr = (C.c_int * len(dts))(*dts)
    ptr_arr = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    rdt = lib().or_eval(expr.encode(), len(arrs), dt_arr, ptr_arr, n, _ptr(out), C.byref(flags))
    if rdt < 0:
        raise ValueError("oracle could not parse " + expr)
    res = out.view(NP_OF[rdt])[:n].copy()
    if rdt == DT_BOOL:
        res = res.astype(np.bool_)
    return res.reshape(inputs[0].shape), flags.value


def sum_seq(x):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    return lib().or_sum_seq_f64(_ptr(x), x.size)


def sum_seq_from(x, start):
    """The sequential Float64 sum continued from the running total `start`."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    return lib().or_sum_seq_f64_from(_ptr(x), x.size, float(start))


def max_seq(x):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    return lib().or_max_seq_f64(_ptr(x), x.size)


def concatenate(arrays, axis=0):
    """NArray.concatenate_to_slice over numpy arrays of one dtype and rank
    (axis >= 0)."""
    arrays = [np.ascontiguousarray(a) for a in arrays]
    rank = arrays[0].ndim
    shape = list(arrays[0].shape)
    shape[axis] = sum(a.shape[axis] for a in arrays)
    out = np.zeros(max(int(np.prod(shape)), 1), dtype=arrays[0].dtype)
    ptrs = (C.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
    shp = _i64([d for a in arrays for d in a.shape] or [0])
    n = lib().or_concatenate(len(arrays), ptrs, shp, rank, arrays[0].itemsize, axis, _ptr(out))
    return out[:n].reshape(shape)


def sum_seq_f32(x):
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
    return lib().or_sum_seq_f32(_ptr(x), x.size)


def batch_layout(shapes):
    """Packed layout of a batch: x/y offsets (16-byte aligned), z offsets,
    totals (the same layout the device batch reports)."""
    xo, zo, ox, oz = [], [], 0, 0
    for s in shapes:
        n = int(np.prod(s))
        xo.append(ox)
        zo.append(oz)
        ox += (n + 3) // 4 * 4
        oz += int(np.prod([(d + 1) // 2 for d in s]))
    return np.array(xo, np.int64), np.array(zo, np.int64), ox, oz


def batch_loop(shapes, scalars, x, xoffs=None, zoffs=None):
    """The per-array host loop (y = x + s, Float32 sequential sum, step-2
    slice) over packed x; returns (y, z, sums)."""
    n = len(shapes)
    lay = batch_layout(shapes)
    if xoffs is None:
        xoffs, zoffs = lay[0], lay[1]
    ranks = np.array([len(s) for s in shapes], np.int32)
    shp = np.zeros((max(n, 1), 3), np.int64)
    for k, s in enumerate(shapes):
        shp[k, :len(s)] = s
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros_like(x)
    z = np.zeros(max(lay[3], 1), np.float32)
    sums = np.zeros(max(n, 1), np.float64)
    sc = np.ascontiguousarray(scalars, dtype=np.float32)
    xoffs = np.ascontiguousarray(xoffs, np.int64)
    zoffs = np.ascontiguousarray(zoffs, np.int64)
    buf = C.create_string_buffer(256)
    _check(lib().or_batch_loop(n, _ptr(ranks), _ptr(shp), _ptr(xoffs), _ptr(zoffs), _ptr(sc), _ptr(x), _ptr(y),
                               _ptr(z), _ptr(sums), buf, 256), buf)
    return y, z[:lay[3]], sums[:n]


def sum_exact(x):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    return lib().or_sum_exact(_ptr(x), x.size)


def sum_i32(x):
    x = np.ascontiguousarray(x, dtype=np.int32).reshape(-1)
    out = C.c_int64(0)
    rc = lib().or_sum_i32(_ptr(x), x.size, C.byref(out))
    return out.value, rc


def slices_reduce(x, axis, mode):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(x.shape[axis], dtype=np.float64)
    lib().or_slices_reduce(_ptr(x), _i64(x.shape), x.ndim, axis, mode, _ptr(out))
    return out


def heat1d_reference(steps=-1):
    out = np.zeros(64, dtype=np.float64)
    n = lib().or_heat1d_reference(steps, _ptr(out))
    return out[:n]


def heat2d(u, r, steps, every=0):
    u = np.ascontiguousarray(u).copy()
    dt = DT_F32 if u.dtype == np.float32 else DT_F64
    ne = steps // every + 1 if every > 0 else 0
    e = np.zeros(max(ne, 1), dtype=np.float64)
    lib().or_heat2d(dt, _ptr(u), u.shape[0], u.shape[1], r, steps, every, _ptr(e) if every > 0 else None)
    return u, e[:ne]


def fill_random(shape, dt, seed, offset=0, dist=0, a=0.0, b=1.0):
    n = int(np.prod(shape))
    out = np.zeros(max(n, 1), dtype=NP_OF[dt])
    lib().or_fill_random(_ptr(out), dt, n, seed, offset, dist, a, b)
    return out[:n].reshape(shape)


def draw(seed, k):
    return lib().or_draw(seed, k)
