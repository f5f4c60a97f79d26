"""Sources, masks and destinations that are views of the array being written
through transform chains longer than one descriptor holds (ComposedTransform
has no bound, src/view_util/transforms.cr:28-101).  The reference reads such
a source element by element as its loop reaches it (src/n_array.cr:484-492
pulls `src.each` lazily; :523-536 evaluates `mask.get(coord)` and
`value.get(coord)` per coordinate), so it sees the loop's earlier writes; the
device reads and writes through the chains' buffer-index arrays
(ph_view_index, ph_map_indexed) in that order.  Checked bit for bit against
the oracle's lazily evaluated loops (or_set_chunk_lazy, or_set_mask_lazy),
and each fixed case also differs from the answer a gathered snapshot would
give, so the order is visible in it.  The C-ABI tests drive ph_view_index
and ph_map_indexed directly, against the loop they state."""
import ctypes
import os
import random
import time

import numpy as np
import pytest

import oracle_api as O
from crystal import parse_literal, trace

pytestmark = pytest.mark.gpu


def _i32(rng, shape, lo=-1000, hi=1000):
    return rng.integers(lo, hi, shape).astype(np.int32)


def _lits(lits):
    return [parse_literal(x) for x in lits]


def _deep(view, shape, chain, rng=None, steps=5):
    """Permutations between reshapes, each reshape a descriptor stage of its
    own, mirrored in the oracle's chain (newest first)."""
    for k in range(steps):
        n = int(np.prod(shape))
        if rng is None:
            ns = [2, n // 2] if k % 2 == 0 else [n // 4, 4]
        else:
            ds = [d for d in range(2, n) if n % d == 0]
            d = rng.choice(ds)
            ns = [d, n // d]
        if len(shape) == 2:
            view = view.permute([1, 0])
            chain.insert(0, ("permute", [1, 0]))
            shape = [shape[1], shape[0]]
        if rng is not None and rng.random() < 0.3:
            view = view.reverse()
            chain.insert(0, ("reverse", list(shape)))
        view = view.reshape(ns)
        chain.insert(0, ("reshape", list(shape), ns))
        shape = ns
    return view, list(shape)


@pytest.mark.parametrize("dtype", [np.int32, np.float64, np.bool_])
def test_deep_source_copied_onto_itself(ph, dtype):
    """A copy moves the elements' bytes as they are, whatever their size."""
    rng = np.random.default_rng(1)
    h = _i32(rng, (24, 16)).astype(dtype) if dtype != np.bool_ else rng.random((24, 16)) < 0.5
    z = ph.NArray.from_numpy(h)
    chain = []
    v, vs = _deep(z.view(), [24, 16], chain)
    assert v.descriptor()["tail"] > 0
    z.reshape(vs)[[]] = v
    ref = h.copy()
    O.set_chunk_lazy(ref.reshape(vs), [], O.region_cover(vs), "x0", [(ref, chain)], vs)
    assert np.array_equal(z.to_numpy(), ref)
    assert not np.array_equal(ref, O.view_gather(h, chain, vs).reshape(24, 16))  # the snapshot answer


@pytest.mark.parametrize("kind", ["i32", "f64"])
def test_deep_procview_source(ph, kind):
    rng = np.random.default_rng(2)
    # a bounded change per link of a dependency chain: no overflow, no inf
    fn = (lambda x: (x ^ 21) - 3) if kind == "i32" else (lambda x: x * 0.5 + 1.25)
    h = _i32(rng, (24, 16)) if kind == "i32" else rng.standard_normal((24, 16))
    z = ph.NArray.from_numpy(h)
    chain = []
    v, vs = _deep(z.view(), [24, 16], chain)
    z.reshape(vs)[[]] = v.process(fn)
    ref = h.copy()
    O.set_chunk_lazy(ref.reshape(vs), [], O.region_cover(vs), trace(fn), [(ref, chain)], vs)
    assert np.array_equal(z.to_numpy(), ref)
    snap, _ = O.eval_expr(trace(fn), [O.view_gather(h, chain, vs)])
    assert not np.array_equal(ref, snap.reshape(24, 16))


def test_mask_and_value_from_deep_views(ph):
    rng = np.random.default_rng(3)
    h = _i32(rng, (24, 16))
    chain = []
    z = ph.NArray.from_numpy(h)
    v, vs = _deep(z.view(), [24, 16], chain)
    zr = z.reshape(vs)
    zr[v.process(lambda x: x > 0)] = 7
    ref = h.copy()
    O.set_mask_lazy(ref.reshape(vs), "(gt x0 0i)", 7, [(ref, chain)])
    assert np.array_equal(z.to_numpy(), ref)
    snap = h.copy().reshape(vs)
    snap[O.view_gather(h, chain, vs) > 0] = 7
    assert not np.array_equal(ref, snap.reshape(24, 16))
    # the value a deep ProcView, the mask a plain array
    fn = lambda x: (x ^ 5) + 1  # noqa: E731
    mh = rng.random(vs) < 0.7
    z = ph.NArray.from_numpy(h)
    v, _ = _deep(z.view(), [24, 16], [])
    z.reshape(vs)[ph.NArray.from_numpy(mh)] = v.process(fn)
    ref = h.copy()
    O.set_mask_lazy(ref.reshape(vs), "x1", trace(fn), [(ref, chain), (mh, [])])
    assert np.array_equal(z.to_numpy(), ref)


def test_deep_destination_with_an_aliased_source(ph):
    rng = np.random.default_rng(4)
    h = _i32(rng, (24, 16))
    # a plain (reversed) view of the same array as the source
    z = ph.NArray.from_numpy(h)
    mchain = []
    mv, ms = _deep(z.mutable_view(), [24, 16], mchain)
    assert mv.descriptor()["tail"] > 0
    mv[[]] = z.reshape(ms).view().reverse()
    ref = h.copy()
    O.set_chunk_lazy(ref, mchain, O.region_cover(ms), "x0", [(ref.reshape(ms), [("reverse", ms)])], ms)
    assert np.array_equal(z.to_numpy(), ref)
    snap = O.view_scatter(h.copy(), mchain, O.region_cover(ms), h.reshape(ms)[::-1, ::-1])
    assert not np.array_equal(ref, snap)
    # a region of the deep destination, from a ProcView of the array itself
    z = ph.NArray.from_numpy(h)
    mv, _ = _deep(z.mutable_view(), [24, 16], [])
    lits = ["1..", "..-3"]
    r = O.region(_lits(lits), ms)
    fn = lambda x: x - 5  # noqa: E731
    mv[lits] = z.reshape(ms).view(*lits).process(fn)
    ref = h.copy()
    O.set_chunk_lazy(ref, mchain, r, trace(fn), [(ref.reshape(ms), [("region", r)])], r.shape)
    assert np.array_equal(z.to_numpy(), ref)


def test_deep_destination_from_a_deep_source(ph):
    rng = np.random.default_rng(5)
    h = _i32(rng, (24, 16))
    z = ph.NArray.from_numpy(h)
    mchain, rchain = [], []
    mv, ms = _deep(z.mutable_view(), [24, 16], mchain)
    rv, rs = _deep(z.view().reverse(), [24, 16], rchain, steps=4)
    rchain.append(("reverse", [24, 16]))
    assert rs != ms
    rv = rv.reshape(ms)
    rchain.insert(0, ("reshape", rs, ms))
    mv[[]] = rv
    ref = h.copy()
    O.set_chunk_lazy(ref, mchain, O.region_cover(ms), "x0", [(ref, rchain)], ms)
    assert np.array_equal(z.to_numpy(), ref)


def test_random_deep_chains_against_the_oracle(ph):
    """Random chains of 3 to 6 reshapes (permutations and reversals between
    them) on either side, copies and checked Int32 maps, with and without a
    deep destination."""
    rng = random.Random(31)
    nrng = np.random.default_rng(31)
    shapes = [[24, 16], [12, 30], [36, 20], [8, 9]]
    for trial in range(24):
        shape = shapes[trial % len(shapes)]
        h = _i32(nrng, shape)
        z = ph.NArray.from_numpy(h)
        rchain, wchain = [], []
        rv, rs = _deep(z.view(), list(shape), rchain, rng=rng, steps=rng.randint(3, 6))
        deep_dst = trial % 3 != 0
        if deep_dst:
            wv, ws = _deep(z.mutable_view(), list(shape), wchain, rng=rng, steps=rng.randint(3, 6))
        else:
            ws = rs
            wv = z.reshape(ws)
        if rs != ws:
            rv = rv.reshape(ws)
            rchain.insert(0, ("reshape", rs, ws))
        use_map = trial % 2 == 1
        fn = lambda x: (x ^ 21) - 3  # noqa: E731
        wv[[]] = rv.process(fn) if use_map else rv
        ref = h.copy()
        expr = trace(fn) if use_map else "x0"
        if deep_dst:
            O.set_chunk_lazy(ref, wchain, O.region_cover(ws), expr, [(ref, rchain)], ws)
        else:
            O.set_chunk_lazy(ref.reshape(ws), [], O.region_cover(ws), expr, [(ref, rchain)], ws)
        assert np.array_equal(z.to_numpy(), ref), (trial, shape, rchain, wchain, use_map)


def test_deep_self_copy_at_256(ph):
    """A 256 x 256 Int32 array copied onto itself through a 5-reshape chain:
    long chains of reads after writes (pointer jumping on the indexed copy),
    and a checked map over the same chain (Jacobi passes, then the loop)."""
    rng = np.random.default_rng(6)
    h = _i32(rng, (256, 256))
    chain = []
    z = ph.NArray.from_numpy(h)
    v, vs = _deep(z.view(), [256, 256], chain)
    z.reshape(vs)[[]] = v
    ref = h.copy()
    O.set_chunk_lazy(ref.reshape(vs), [], O.region_cover(vs), "x0", [(ref, chain)], vs)
    assert np.array_equal(z.to_numpy(), ref)
    z = ph.NArray.from_numpy(h)
    v, _ = _deep(z.view(), [256, 256], [])
    fn = lambda x: x // 3 + 1  # noqa: E731
    z.reshape(vs)[[]] = v.process(fn)
    ref = h.copy()
    O.set_chunk_lazy(ref.reshape(vs), [], O.region_cover(vs), trace(fn), [(ref, chain)], vs)
    assert np.array_equal(z.to_numpy(), ref)


def test_deep_overflow_is_raised(ph):
    """A checked Int32 map through a deep aliased source raises as the
    reference's `*` would."""
    h = np.full((24, 16), 2 ** 30, np.int32)
    z = ph.NArray.from_numpy(h)
    v, vs = _deep(z.view(), [24, 16], [])
    with pytest.raises(OverflowError):
        z.reshape(vs)[[]] = v.process(lambda x: x * 4)


# --- the C-ABI -------------------------------------------------------------

class _Affine(ctypes.Structure):
    _fields_ = [("rank", ctypes.c_int32), ("pad_", ctypes.c_int32), ("offset", ctypes.c_int64),
                ("shape", ctypes.c_int64 * 8), ("stride", ctypes.c_int64 * 8)]


class _Desc(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("nstage", ctypes.c_int32), ("extent", ctypes.c_int64),
                ("view", _Affine), ("stage", _Affine * 2)]


def _desc(dtype, extent, shape, strides=None, offset=0):
    d = _Desc()
    d.dtype, d.nstage, d.extent = dtype, 0, extent
    d.view.rank, d.view.offset = len(shape), offset
    st = strides or [int(np.prod(shape[i + 1:])) for i in range(len(shape))]
    for i, (n, s) in enumerate(zip(shape, st)):
        d.view.shape[i], d.view.stride[i] = n, s
    return d


def _lib(ph):
    lib = ctypes.CDLL(os.path.join(ph.LIB_DIR, "libphhip.so"))
    lib.ph_view_index.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Desc)]
    lib.ph_map_indexed.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p),
                                   ctypes.POINTER(_Desc), ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                   ctypes.POINTER(_Desc), ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
    lib.ph_last_error.restype = ctypes.c_char_p
    return lib


def _dev_i64(ph, a):
    """A device array holding int64 values (as 8-byte elements)."""
    return ph.NArray.from_numpy(np.ascontiguousarray(a, np.int64).view(np.float64))


def _copy_indexed(lib, buf, n, ri, wi):
    ins = (ctypes.c_void_p * 1)(buf.device_ptr)
    ix = (ctypes.c_void_p * 1)(ri.device_ptr if ri is not None else None)
    sd = _desc(0, buf.size, [n])
    od = _desc(0, buf.size, [n])
    fl = ctypes.c_int32(0)
    return lib.ph_map_indexed(None, 1, ins, ctypes.byref(sd), ix, buf.device_ptr, ctypes.byref(od),
                              wi.device_ptr if wi is not None else None, ctypes.byref(fl))


def test_view_index_of_a_staged_descriptor(ph):
    """ph_view_index of a reshaped, permuted view (one stage) equals the
    buffer indices numpy gives for the same view of an arange."""
    lib = _lib(ph)
    d = _desc(0, 384, [16, 24], strides=[1, 16])  # the transpose of a 24 x 16 array
    d.nstage = 1
    st = d.stage[0]
    st.rank, st.offset = 2, 0
    st.shape[0], st.shape[1] = 24, 16
    st.stride[0], st.stride[1] = 1, 24  # then read transposed again as 16 x 24 storage
    out = ph.NArray.empty([384], ph.DType.F64)
    assert lib.ph_view_index(out.device_ptr, ctypes.byref(d)) == 0, lib.ph_last_error()
    got = out.to_numpy().view(np.int64)
    lin = np.arange(384).reshape(24, 16).T.reshape(-1)  # affine part: transposed 24 x 16
    base = np.arange(384).reshape(16, 24).T.reshape(-1)  # stage: 24 x 16 coordinates into 16 x 24 storage
    assert np.array_equal(got, base[lin])


def test_map_indexed_copy_follows_the_loop(ph):
    """buf[w[k]] = buf[r[k]] for k in order, with random index arrays over
    one Int32 buffer: the device result is the sequential loop's."""
    lib = _lib(ph)
    rng = np.random.default_rng(7)
    for n, m in [(500, 700), (4096, 4096), (10000, 10000)]:
        h = _i32(rng, (m,))
        w = rng.permutation(m)[:n]
        r = rng.integers(0, m, n)
        buf = ph.NArray.from_numpy(h)
        rc = _copy_indexed(lib, buf, n, _dev_i64(ph, r), _dev_i64(ph, w))
        assert rc == 0, lib.ph_last_error()
        ref = h.copy()
        for k in range(n):
            ref[w[k]] = ref[r[k]]
        assert np.array_equal(buf.to_numpy(), ref), (n, m)


def test_map_indexed_rejects_bad_indexes(ph):
    lib = _lib(ph)
    h = np.arange(100, dtype=np.int32)
    buf = ph.NArray.from_numpy(h)
    r = np.arange(100)
    r[37] = 100  # past the buffer
    rc = _copy_indexed(lib, buf, 100, _dev_i64(ph, r), _dev_i64(ph, np.arange(100)[::-1]))
    assert rc == -1 and b"outside" in lib.ph_last_error()
    w = np.arange(100)
    w[5] = 6  # element 6 named twice
    rc = _copy_indexed(lib, buf, 100, _dev_i64(ph, np.arange(100)), _dev_i64(ph, w))
    assert rc == -4 and b"twice" in lib.ph_last_error()
    assert np.array_equal(buf.to_numpy(), h)  # nothing written


class _Insn(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint8), ("dt", ctypes.c_uint8), ("dst", ctypes.c_uint8), ("a", ctypes.c_uint8),
                ("b", ctypes.c_uint8), ("c", ctypes.c_uint8), ("pad_", ctypes.c_uint16), ("imm", ctypes.c_int64)]


class _Prog(ctypes.Structure):
    _fields_ = [("n_insn", ctypes.c_int32), ("n_in", ctypes.c_int32), ("out_reg", ctypes.c_int32),
                ("out_dtype", ctypes.c_int32), ("insn", _Insn * 48)]


def _prog(ph, e, n_in):
    """A ph_prog from the host's lowering of expression `e`."""
    insns, out_reg, out_dt = ph.program(e, n_in)
    p = _Prog()
    p.n_insn, p.n_in, p.out_reg, p.out_dtype = len(insns), n_in, out_reg, int(out_dt)
    for k, (op, dt, dst, a, b, c, imm) in enumerate(insns):
        I = p.insn[k]
        I.op, I.dt, I.dst, I.a, I.b, I.c = op, dt, dst, a, b, c
        I.imm = (np.array([imm], np.float64).view(np.int64)[0] if isinstance(imm, float) else int(imm))
    return p


def test_map_indexed_programs_follow_the_loop(ph):
    """ph_map_indexed with a program, input 0 read from the output's buffer
    through an index array: buf[w[k]] = f(buf[r[k]], y[k]) for k in order.
    One chain of 6000 dependent elements (a shift by one), 64 interleaved
    chains (a shift by 64: a thread per chain), the same with one element
    read by two later ones (a tree whose inner elements still form chains),
    a random tree (each element reads one of the 63 written before it: a
    pass per depth level), a tree as deep as the chain (the in-order loop),
    and a second input y read through a plain descriptor from another
    buffer."""
    lib = _lib(ph)
    lib.ph_map_indexed.argtypes = [ctypes.POINTER(_Prog), ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p),
                                   ctypes.POINTER(_Desc), ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                   ctypes.POINTER(_Desc), ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
    I = ph.DType.I32
    e = ((ph.input(0, I) ^ 21) - 3) + ph.input(1, I)
    prog = _prog(ph, e, 2)
    rng = np.random.default_rng(11)
    n = 6000
    cases = {"one chain": np.r_[0, np.arange(n - 1)], "64 chains": np.r_[np.arange(64), np.arange(n - 64)]}
    tree = cases["64 chains"].copy()
    tree[200] = 72  # element 72 is read by 136 and by 200
    cases["tree"] = tree
    rt = np.arange(n) - rng.integers(1, 64, n)
    cases["random tree"] = np.where(rt < 0, np.arange(n), rt)
    deep = cases["one chain"].copy()
    deep[3000] = 17  # element 17 is read by 18 and by 3000
    cases["deep tree"] = deep
    for name, r in cases.items():
        print(f"indexed program: {name}", flush=True)
        h = rng.integers(-1000, 1000, n).astype(np.int32)
        yh = rng.integers(-3, 4, n).astype(np.int32)
        w = np.arange(n)
        buf = ph.NArray.from_numpy(h)
        y = ph.NArray.from_numpy(yh)
        ri, wi = _dev_i64(ph, r), _dev_i64(ph, w)
        ins = (ctypes.c_void_p * 2)(buf.device_ptr, y.device_ptr)
        ds = (_Desc * 2)(_desc(0, n, [n]), _desc(0, n, [n]))
        ix = (ctypes.c_void_p * 2)(ri.device_ptr, None)
        od = _desc(0, n, [n])
        fl = ctypes.c_int32(0)
        rc = lib.ph_map_indexed(ctypes.byref(prog), 2, ins, ds, ix, buf.device_ptr, ctypes.byref(od), wi.device_ptr,
                                ctypes.byref(fl))
        assert rc == 0 and fl.value == 0, lib.ph_last_error()
        ref = h.astype(np.int64)
        for k in range(n):
            ref[w[k]] = ((ref[r[k]] ^ 21) - 3) + yh[k]
        assert np.array_equal(buf.to_numpy(), ref.astype(np.int32)), name


def test_map_indexed_random_tree_by_levels(ph):
    """A random tree of 200000 elements, each reading one of the 63 written
    just before it (about 6000 depth levels, so a pass per level), with the
    output written through a permutation: the device result is the loop's,
    buf[w[k]] = f(buf[r[k]], y[k]) for k in order."""
    lib = _lib(ph)
    lib.ph_map_indexed.argtypes = [ctypes.POINTER(_Prog), ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p),
                                   ctypes.POINTER(_Desc), ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                   ctypes.POINTER(_Desc), ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
    I = ph.DType.I32
    prog = _prog(ph, ((ph.input(0, I) ^ 21) - 3) + ph.input(1, I), 2)
    rng = np.random.default_rng(12)
    n = 200000
    w = rng.permutation(n)
    back = rng.integers(1, 64, n)
    k = np.arange(n)
    r = np.where(k - back >= 0, w[np.maximum(k - back, 0)], rng.integers(0, n, n))  # read what element k - back wrote
    h = rng.integers(-1000, 1000, n).astype(np.int32)
    yh = rng.integers(-3, 4, n).astype(np.int32)
    buf, y = ph.NArray.from_numpy(h), ph.NArray.from_numpy(yh)
    ri, wi = _dev_i64(ph, r), _dev_i64(ph, w)
    ins = (ctypes.c_void_p * 2)(buf.device_ptr, y.device_ptr)
    ds = (_Desc * 2)(_desc(0, n, [n]), _desc(0, n, [n]))
    ix = (ctypes.c_void_p * 2)(ri.device_ptr, None)
    od = _desc(0, n, [n])
    fl = ctypes.c_int32(0)
    ph.sync()
    t0 = time.perf_counter()
    rc = lib.ph_map_indexed(ctypes.byref(prog), 2, ins, ds, ix, buf.device_ptr, ctypes.byref(od), wi.device_ptr,
                            ctypes.byref(fl))
    ph.sync()
    dt = time.perf_counter() - t0
    assert rc == 0 and fl.value == 0, lib.ph_last_error()
    ref = h.astype(np.int64)
    rl, wl, yl = r.tolist(), w.tolist(), yh.tolist()
    refl = ref.tolist()
    for j in range(n):
        refl[wl[j]] = ((refl[rl[j]] ^ 21) - 3) + yl[j]
    assert np.array_equal(buf.to_numpy(), np.array(refl, np.int64).astype(np.int32))
    print(f"random tree of {n} elements by levels: {dt * 1e3:.1f} ms with compilation")
