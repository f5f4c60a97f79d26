"""Parity of the device path (HIP kernels through the C-ABI, driven by the
host API) with the CPU oracle, on the same seeded inputs.

Integer / index work is compared bit for bit; Float64 elementwise work too
(IEEE-754 correct rounding, no contraction); reductions against the oracle's
correctly rounded sums with the tolerance stated per test."""
import json
import os
import random

import numpy as np
import pytest

import oracle_api as O
from crystal import parse_literal, trace

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_spec_vectors.json")))


def P(lits):
    return [parse_literal(x) for x in lits]


def _lit(x):
    return x


# ------------------------------------------------------------- known answers
def test_spec_chunks_on_device(ph):
    g = GOLD["narray_chunks"]
    stock = np.array(g["stock"], np.int32).reshape(g["stock_shape"])
    d = ph.NArray.from_numpy(stock)
    for lit, exp in g["fetch"]:
        assert d[lit].to_numpy().reshape(-1).tolist() == exp
    assert d[g["fetch_empty"]].size == 0
    for lit, src, exp in g["set_src"]:
        e = d.clone()
        e[lit] = ph.NArray.from_numpy(np.array(src, np.int32))
        assert e.to_numpy().reshape(-1).tolist() == exp
    for lit, v, exp in g["set_scalar"]:
        e = d.clone()
        e[lit] = v
        assert e.to_numpy().reshape(-1).tolist() == exp
    m = GOLD["narray_masks"]
    mask = ph.NArray.from_numpy(np.array(m["mask"]).reshape(2, 3))
    e = d.clone()
    e[mask] = m["scalar"]
    assert e.to_numpy().reshape(-1).tolist() == m["expected_scalar"]
    e = d.clone()
    e[mask] = d + m["src_plus"]
    assert e.to_numpy().reshape(-1).tolist() == m["expected_src"]
    mp = GOLD["narray_maps"]
    assert d.map(lambda x: x ** 2).to_numpy().reshape(-1).tolist() == mp["squared"]
    assert (d * 3).to_numpy().reshape(-1).tolist() == mp["times3"]
    e = d.clone()
    e.map_(lambda x: x ** 2)
    assert e.to_numpy().reshape(-1).tolist() == mp["squared"]
    w = GOLD["multi_writable"]
    buf = ph.NArray.from_numpy(np.array(w["buffer"], np.int32).reshape(w["shape"]))
    b2 = buf.clone()
    b2[w["region"]] = w["scalar"]
    assert b2.to_numpy().reshape(-1).tolist() == w["expected_scalar"]
    b2 = buf.clone()
    b2[w["region"]] = ph.NArray.from_numpy(np.array(w["chunk"], np.int32).reshape(w["chunk_shape"]))
    assert b2.to_numpy().reshape(-1).tolist() == w["expected_chunk"]
    with pytest.raises(ph.ShapeError):
        b2[w["bad_region"]] = ph.NArray.from_numpy(np.array(w["chunk"], np.int32).reshape(w["chunk_shape"]))
    it = GOLD["iteration_order"]
    a = ph.NArray.from_numpy(np.array(it["values"], np.int32).reshape(it["shape"]))
    assert a.permute(1, 0).to_numpy().reshape(-1).tolist() == it["colex"]


def test_transform_known_answers_on_views(ph):
    t = GOLD["transforms"]
    rs = t["reshape"]
    src = ph.NArray.from_numpy(np.arange(12, dtype=np.int32).reshape(rs["src_shape"]))
    v = src.view().reshape(rs["new_shape"])
    assert v.apply(rs["coord"]) == rs["expected"]
    assert v.get(*rs["coord"]) == int(np.arange(12).reshape(rs["src_shape"])[tuple(rs["expected"])])
    rv = t["reverse"]
    src = ph.NArray.from_numpy(np.arange(15, dtype=np.int32).reshape(rv["shape"]))
    assert src.view().reverse().apply(rv["coord"]) == rv["expected"]
    pm = t["permute"]
    shape = [max(pm["coord"][i] + 1, 1) for i in range(7)]
    # the view axis i reads source axis pattern[i]
    src_shape = [0] * 7
    for i, p in enumerate(pm["pattern"]):
        src_shape[p] = shape[i]
    src = ph.NArray.fill(src_shape, 0)
    assert src.view().permute(pm["pattern"]).apply(pm["coord"]) == pm["expected"]


# ----------------------------------------------------------- gather / scatter
def _random_literal(rng, bound):
    kind = rng.random()
    if kind < 0.15:
        return rng.randint(-bound, bound - 1) if bound else 0
    f = rng.choice([None, rng.randint(-bound, max(bound - 1, 0))])
    l = rng.choice([None, rng.randint(-bound, max(bound - 1, 0))])
    s = rng.choice([None, None, rng.choice([-3, -2, -1, 1, 2, 3])])
    ex = rng.random() < 0.3
    fs = "" if f is None else str(f)
    ls = "" if l is None else str(l)
    dots = "..." if ex else ".."
    if s is None:
        return f"{fs}{dots}{ls}"
    return f"{fs}..{s}{dots}{ls}"


@pytest.mark.parametrize("dtype", ["I32", "F64", "BF16", "Bool"])
def test_random_gather_scatter(ph, dtype):
    rng = random.Random(hash(dtype) % 1000)
    D = getattr(ph.DType, dtype)
    npdt = {"I32": np.int32, "F64": np.float64, "BF16": np.uint16, "Bool": np.bool_}[dtype]
    done = 0
    for trial in range(120):
        rank = rng.randint(1, 4)
        shape = [rng.randint(1, 9) for _ in range(rank)]
        lits = [_random_literal(rng, b) for b in shape[: rng.randint(1, rank)]]
        try:
            r = O.region(P(lits), shape)
        except O.OracleError:
            with pytest.raises((ph.IndexError, ph.ShapeError, ph.DivisionByZeroError)):
                ph.IndexRegion(lits, shape)
            continue
        host = np.random.default_rng(trial).integers(0, 2 ** 15, size=shape).astype(npdt)
        dev = ph.NArray.from_numpy(host, D)
        got = dev[lits]
        exp = O.gather(host, r)
        assert got.shape == r.shape, (shape, lits)
        assert np.array_equal(got.to_numpy().reshape(-1), exp.reshape(-1)), (shape, lits)
        # view(r).to_narr must equal narr[r]
        v = dev.view(*lits)
        assert np.array_equal(v.to_narr().to_numpy().reshape(-1), exp.reshape(-1))
        # scatter back into zeros through a MutableView (src in lex order)
        if r.size:
            src = (np.arange(r.size) % 251 + 1).astype(npdt).reshape(r.shape)
            z = ph.NArray.fill(shape, 0, D)
            z.mutable_view(*lits)[[]] = ph.NArray.from_numpy(src, D)
            ref = O.scatter(np.zeros(shape, npdt), r, src)
            assert np.array_equal(z.to_numpy(), ref), (shape, lits)
            z2 = ph.NArray.from_numpy(np.ones(shape, npdt), D)
            z2[lits] = ph.NArray.from_numpy(src, D)
            ref2 = O.scatter(np.ones(shape, npdt), r, src)
            assert np.array_equal(z2.to_numpy(), ref2), (shape, lits)
        done += 1
    assert done > 40


def _random_chain(rng, view, vshape, chain):
    for _ in range(rng.randint(1, 4)):
        op = rng.choice(["region", "permute", "reverse", "reshape"])
        if op == "region":
            lits = [_random_literal(rng, b) for b in vshape]
            try:
                r = O.region(P(lits), vshape)
            except O.OracleError:
                continue
            if r.size == 0:
                continue
            view = view.mutable_view(*lits) if hasattr(view, "mutable_view") else view.view(*lits)
            chain.insert(0, ("region", r))
            vshape = r.shape
        elif op == "permute":
            pat = list(range(len(vshape)))
            rng.shuffle(pat)
            view = view.permute(pat)
            chain.insert(0, ("permute", pat))
            vshape = [vshape[p] for p in pat]
        elif op == "reverse":
            view = view.reverse()
            chain.insert(0, ("reverse", list(vshape)))
        else:
            n = int(np.prod(vshape))
            divs = [d for d in range(1, n + 1) if n % d == 0]
            d = rng.choice(divs)
            ns = [d, n // d] if rng.random() < 0.5 else [n]
            view = view.reshape(ns)
            chain.insert(0, ("reshape", list(vshape), ns))
            vshape = ns
    return view, vshape


def test_view_chains_and_mutable_views(ph):
    rng = random.Random(17)
    for trial in range(80):
        rank = rng.randint(1, 3)
        shape = [rng.randint(1, 6) for _ in range(rank)]
        host = np.arange(int(np.prod(shape)), dtype=np.int32).reshape(shape) * 3 + 1
        dev = ph.NArray.from_numpy(host)
        chain = []
        view, vshape = _random_chain(rng, dev.view(), list(shape), chain)
        exp = O.view_gather(host, chain, vshape)
        assert list(view.shape) == list(vshape)
        assert np.array_equal(view.to_narr().to_numpy(), exp), chain
        # elementwise through the view (the view is an operand of a fused map)
        assert np.array_equal((view + 1).to_numpy(), exp + 1)
        # write-through: MutableView with the same chain
        if not chain:
            continue
        mv = dev.clone().mutable_view()
        mchain = []
        mrng = random.Random(trial)
        mview, mshape = _random_chain(mrng, mv, list(shape), mchain)
        if not isinstance(mview, ph.MutableView) or int(np.prod(mshape)) == 0:
            continue
        src = (np.arange(int(np.prod(mshape)), dtype=np.int32) + 1000).reshape(mshape)
        target = ph.NArray.from_numpy(host.copy())
        tv = target.mutable_view()
        mrng = random.Random(trial)
        tv2, _ = _random_chain(mrng, tv, list(shape), [])
        tv2[[]] = ph.NArray.from_numpy(src)
        reg = O.region_cover(mshape)
        ref = O.view_scatter(host.copy(), mchain, reg, src)
        assert np.array_equal(target.to_numpy(), ref), mchain


def _deep_chain(view, shape, chain):
    """Reshapes between permutations, each needing a descriptor stage of its
    own: more than the descriptor holds."""
    for k in range(5):
        n = int(np.prod(shape))
        ns = [2, n // 2] if k % 2 == 0 else [n // 4, 4]
        view = view.permute([1, 0]) if len(shape) == 2 else view
        if len(shape) == 2:
            chain.insert(0, ("permute", [1, 0]))
            shape = [shape[1], shape[0]]
        view = view.reshape(ns)
        chain.insert(0, ("reshape", list(shape), ns))
        shape = ns
    return view, shape


def test_transform_chains_past_the_descriptor(ph):
    """ComposedTransform has no bound (src/view_util/transforms.cr:28-101):
    a chain with more reshape stages than one descriptor holds is read (the
    intermediate gathered when used, so later writes to the source are seen)
    and written through (MutableView), bit-exact against the oracle."""
    host = (np.arange(24 * 16, dtype=np.int32) * 7 - 5).reshape(24, 16)
    dev = ph.NArray.from_numpy(host)
    chain = []
    view, vshape = _deep_chain(dev.view(), [24, 16], chain)
    assert view.descriptor()["tail"] > 0
    assert np.array_equal(view.to_narr().to_numpy(), O.view_gather(host, chain, vshape))
    assert np.array_equal((view * 3).to_numpy(), O.view_gather(host, chain, vshape) * 3)
    dev[["0..3", ".."]] = 11  # the view is lazy: it reads the source when used
    h2 = host.copy()
    h2[0:4] = 11
    assert np.array_equal(view.to_narr().to_numpy(), O.view_gather(h2, chain, vshape))
    assert view.get([1, 2]) == int(O.view_gather(h2, chain, vshape)[1, 2])
    # write through
    target = ph.NArray.from_numpy(host.copy())
    mchain = []
    mview, mshape = _deep_chain(target.mutable_view(), [24, 16], mchain)
    assert mview.descriptor()["tail"] > 0
    src = (np.arange(int(np.prod(mshape)), dtype=np.int32) + 1000).reshape(mshape)
    mview[[]] = ph.NArray.from_numpy(src)
    assert np.array_equal(target.to_numpy(), O.view_scatter(host.copy(), mchain, O.region_cover(mshape), src))
    target = ph.NArray.from_numpy(host.copy())
    mview, _ = _deep_chain(target.mutable_view(), [24, 16], [])
    mview[["1..2..", "1"]] = -3
    sub = O.region(P(["1..2..", "1"]), mshape)
    exp = O.view_scatter(host.copy(), mchain, sub, np.full(sub.size, -3, np.int32))
    assert np.array_equal(target.to_numpy(), exp)
    mview.set_element([0, 0], 77)
    exp2 = O.view_scatter(exp.copy(), mchain, O.region(P(["0", "0"]), mshape, drop=False), np.int32([77]))
    assert np.array_equal(target.to_numpy(), exp2)


def test_tile_views(ph):
    rng = np.random.default_rng(9)
    for shape, counts in [([1, 5], [4, 1]), ([2, 3], [2, 3]), ([3], [5]), ([2, 1, 2], [1, 3, 2]), ([1, 1], [3, 4])]:
        host = rng.integers(-100, 100, size=shape).astype(np.int32)
        dev = ph.NArray.from_numpy(host)
        t = dev.tile(*counts)
        assert np.array_equal(t.to_narr().to_numpy(), O.tile(host, counts))
        assert np.array_equal((t * 2).to_numpy(), O.tile(host, counts) * 2)


def test_aliased_scatter_follows_lex_order(ph):
    """narr[1..] = narr.view(..-2): ph-core reads each source element after the
    writes before it (a propagating copy)."""
    host = np.arange(10, dtype=np.int32)
    dev = ph.NArray.from_numpy(host)
    dev[["1.."]] = dev.view("..-2")
    ref = host.copy()
    for k in range(9):  # sequential lex-order semantics
        ref[1 + k] = ref[k]
    assert dev.to_numpy().tolist() == ref.tolist()


def test_errors_match_reference(ph):
    d = ph.NArray.fill([3, 4], 0)
    with pytest.raises(ph.IndexError):
        d[["0..5"]]
    with pytest.raises(ph.DimensionError):
        d[[0, 0, 0]]
    with pytest.raises(ph.ShapeError):
        d + ph.NArray.fill([4, 3], 0)
    with pytest.raises(ph.IndexError):
        d.view().permute([0, 2])
    with pytest.raises(ph.ShapeError):
        d.view().reshape([5, 2])
    with pytest.raises(ph.EmptyError):
        d[["0...0"]].max()


# ----------------------------------------------------------------- elementwise
INT_EXPRS = [
    lambda a, b: a + b, lambda a, b: a - b, lambda a, b: a * b, lambda a, b: a // b,
    lambda a, b: a % b, lambda a, b: a / b, lambda a, b: a.wadd(b), lambda a, b: a.wmul(b),
    lambda a, b: (a & b) ^ (a | 7), lambda a, b: a > b, lambda a, b: a.cmp(b), lambda a, b: -a + ~b,
    lambda a, b: (a + 3) * 2 - b // 5, lambda a, b: a ** 2,
]
FLOAT_EXPRS = [
    lambda a, b: a + b, lambda a, b: a - b, lambda a, b: a * b, lambda a, b: a / b,
    lambda a, b: (a + 4) ** 2 * 0.5 - a, lambda a, b: a // b, lambda a, b: a % b,
    lambda a, b: (a * b - a) * (b + 1.5), lambda a, b: a ** 3, lambda a, b: a >= b,
    lambda a, b: -a * 2 + b / 3, lambda a, b: a.abs() - b,
]


@pytest.mark.parametrize("k", range(len(INT_EXPRS)))
def test_int_elementwise_bit_exact(ph, k):
    fn = INT_EXPRS[k]
    rng = np.random.default_rng(k)
    a = rng.integers(-2000, 2000, size=(37, 53)).astype(np.int32)
    b = rng.integers(-40, 40, size=(37, 53)).astype(np.int32)
    b[b == 0] = 3
    exp, flags = O.eval_expr(trace(fn, 2), [a, b])
    assert flags == 0
    got = ph.NArray.from_numpy(a).map_with(ph.NArray.from_numpy(b), block=fn).to_numpy()
    assert got.dtype == exp.dtype and np.array_equal(got, exp)


@pytest.mark.parametrize("k", range(len(FLOAT_EXPRS)))
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_float_elementwise_bit_exact(ph, k, dt):
    fn = FLOAT_EXPRS[k]
    rng = np.random.default_rng(100 + k)
    a = rng.standard_normal((41, 67)).astype(dt)
    b = (rng.standard_normal((41, 67)) + 3.0).astype(dt)
    exp, _ = O.eval_expr(trace(fn, 2), [a, b])
    got = ph.NArray.from_numpy(a).map_with(ph.NArray.from_numpy(b), block=fn).to_numpy()
    assert got.dtype == exp.dtype
    assert np.array_equal(got, exp)


def test_int32_overflow_and_division_errors(ph):
    a = ph.NArray.from_numpy(np.array([2 ** 31 - 1, 5], np.int32))
    with pytest.raises(ph.OverflowError):
        a + 1
    assert (a.wadd(1)).to_numpy().tolist() == [-2 ** 31, 6]
    with pytest.raises(ph.DivisionByZeroError):
        a // ph.NArray.from_numpy(np.array([1, 0], np.int32))
    with pytest.raises(ph.ArgumentError):
        a ** -1
    with pytest.raises(ph.OverflowError):
        ph.NArray.from_numpy(np.array([2 ** 30, 2 ** 30], np.int32)).sum()


def test_scalar_first_and_procview_fusion(ph):
    x = np.random.default_rng(5).standard_normal((33, 65)).astype(np.float32)
    d = ph.NArray.from_numpy(x)
    got = (2 * d).to_numpy()  # patches/number.cr: Int32 * Float32 -> Float32
    assert got.dtype == np.float32 and np.array_equal(got, np.float32(2) * x)
    pv = d.process(lambda v: (v + 4) ** 2 * 0.5 - v)
    exp, _ = O.eval_expr(trace(lambda v: (v + 4) ** 2 * 0.5 - v), [x])
    assert np.array_equal(pv.to_narr().to_numpy(), exp)
    # a ProcView operand fused into a binary op, and a chained block
    exp2, _ = O.eval_expr(trace(lambda v, w: ((v + 4) ** 2 * 0.5 - v) * w, 2), [x, x])
    assert np.array_equal((pv * d).to_numpy(), exp2)
    exp3, _ = O.eval_expr(trace(lambda v: ((v + 4) ** 2 * 0.5 - v) - 1.0), [x])
    assert np.array_equal(pv.process(lambda y: y - 1.0).to_narr().to_numpy(), exp3)
    # a block over a strided view
    v = d.view("1..2..", "..-1..").process(lambda t: t * t)
    xs = x[1::2, ::-1]
    assert np.array_equal(v.to_narr().to_numpy(), xs * xs)


# ------------------------------------------------------- C1: broadcast FP64
def test_broadcast_config_c1(ph):
    """(a + b.tile(1024, 1)) * c, Float64, U[-1,1) from seed 0 (streams a, b,
    c in that order); bit-exact against the oracle."""
    n = 1024
    a = ph.NArray.random([n, n], ph.DType.F64, seed=0, offset=0, a=-1.0, b=1.0)
    b = ph.NArray.random([1, n], ph.DType.F64, seed=0, offset=n * n, a=-1.0, b=1.0)
    c = ph.NArray.random([n, n], ph.DType.F64, seed=0, offset=n * n + n, a=-1.0, b=1.0)
    an = O.fill_random([n, n], O.DT_F64, 0, 0, 0, -1.0, 1.0)
    bn = O.fill_random([1, n], O.DT_F64, 0, n * n, 0, -1.0, 1.0)
    cn = O.fill_random([n, n], O.DT_F64, 0, n * n + n, 0, -1.0, 1.0)
    assert np.array_equal(a.to_numpy(), an) and np.array_equal(b.to_numpy(), bn)
    got = ((a + b.tile(n, 1)) * c).to_numpy()
    exp, _ = O.eval_expr("(mul (add x0 x1) x2)", [an, O.tile(bn, [n, 1]), cn])
    assert np.array_equal(got, exp)


# -------------------------------------------------- C2: strided gather/scatter
def test_gather_scatter_config_c2(ph):
    """Int32 8192x8192 U[0,2^31) seed 1; view(1..3..-1, ..-2..) gathered to a
    contiguous array, then written back into a zeroed array through a
    MutableView (fused fill + scatter).  Bit-exact."""
    n = 8192
    x = ph.NArray.random([n, n], ph.DType.I32, seed=1, dist="int31")
    xn = x.to_numpy()
    assert np.array_equal(xn[:2].reshape(-1), O.fill_random([2 * n], O.DT_I32, 1, 0, 2))
    lits = ["1..3..-1", "..-2.."]
    g = x.view(*lits).to_narr()
    assert g.shape == [2731, 4096]
    r = O.region(P(lits), [n, n])
    exp = O.gather(xn, r)
    gn = g.to_numpy()
    assert np.array_equal(gn, exp)
    z = ph.NArray.fill([n, n], 0)
    z.mutable_view(*lits)[[]] = g
    zn = z.to_numpy()
    ref = np.zeros_like(xn)
    ref[1::3, ::-2] = xn[1::3, ::-2]
    assert np.array_equal(zn, ref)
    # the same write-back without the fused path (array already materialised)
    z2 = ph.NArray.fill([n, n], 0)
    z2.get(0, 0)
    z2.mutable_view(*lits)[[]] = g
    assert np.array_equal(z2.to_numpy(), ref)


def test_device_random_matches_oracle(ph):
    for dt, dist, a, b in [(ph.DType.F64, "uniform", -1.0, 1.0), (ph.DType.F32, "uniform", -1.0, 1.0),
                           (ph.DType.I32, "int31", 0, 0), (ph.DType.I32, "uniform", -5, 7)]:
        d = ph.NArray.random([1000], dt, seed=3, offset=17, dist=dist, a=a, b=b).to_numpy()
        odt = {ph.DType.F64: O.DT_F64, ph.DType.F32: O.DT_F32, ph.DType.I32: O.DT_I32}[dt]
        o = O.fill_random([1000], odt, 3, 17, {"uniform": 0, "normal": 1, "int31": 2}[dist], a, b)
        assert np.array_equal(d, o)
    d = ph.NArray.random([1000], ph.DType.F64, seed=2, dist="normal", a=0.0, b=1.0).to_numpy()
    o = O.fill_random([1000], O.DT_F64, 2, 0, 1, 0.0, 1.0)
    assert np.max(np.abs(d - o)) < 1e-13


# ------------------------------------------ C3: ProcView + reductions FP32/BF16
def _c3_block(x32):
    """The block evaluated by the oracle with Crystal's promotion rules:
    (x+4)**2 in Float32, then * 0.5 and - x in Float64."""
    f, _ = O.eval_expr(trace(lambda v: (v + 4) ** 2 * 0.5 - v), [x32])
    return f


@pytest.mark.parametrize("storage", ["F32", "BF16"])
def test_reductions_small(ph, storage):
    rows, cols = 1000, 1536
    D = getattr(ph.DType, storage)
    x = ph.NArray.random([rows, cols], D, seed=2, dist="normal", a=0.0, b=1.0)
    if storage == "F32":
        x32 = x.to_numpy()
    else:
        x32 = (x.to_numpy().astype(np.uint32) << 16).view(np.float32)
    pv = x.process(lambda v: (v + 4) ** 2 * 0.5 - v)
    res = pv.reduce2d_all()
    f = _c3_block(x32)
    full = O.sum_exact(f)
    assert abs(res["full_sum"] - full) <= 1e-12 * abs(full)
    assert res["full_max"] == float(f.max())  # max is order-free: exact
    rs = np.array([O.sum_exact(f[i]) for i in range(rows)])
    cs = np.array([O.sum_exact(f[:, j]) for j in range(cols)])
    assert np.max(np.abs(np.array(res["row_sum"]) - rs) / np.abs(rs)) < 1e-12
    assert np.max(np.abs(np.array(res["col_sum"]) - cs) / np.abs(cs)) < 1e-12
    assert np.array_equal(res["row_max"], f.max(axis=1))
    assert np.array_equal(res["col_max"], f.max(axis=0))
    # slices(k).map(&.sum) and the full sum through the generic entry point
    assert np.allclose(pv.slices_sum(0), rs, rtol=1e-12, atol=0)
    assert np.allclose(pv.slices_sum(1), cs, rtol=1e-12, atol=0)
    assert pv.sum() == pytest.approx(full, rel=1e-12)
    assert pv.sum(acc="f64") == pytest.approx(full, rel=1e-12)
    # max is exact (order-free)
    assert pv.max() == float(np.max(f))


def test_integer_sums_exact(ph):
    rng = np.random.default_rng(8)
    a = rng.integers(-1000, 1000, size=(300, 700)).astype(np.int32)
    d = ph.NArray.from_numpy(a)
    assert d.sum() == int(a.sum())
    assert d.max() == int(a.max())
    assert np.array_equal(np.array(d.slices_sum(0)), a.sum(axis=1).astype(np.float64))
    assert np.array_equal(np.array(d.slices_sum(1)), a.sum(axis=0).astype(np.float64))
    assert np.array_equal(np.array(d.slices_max(1)), a.max(axis=0).astype(np.float64))
    b = rng.integers(-100, 100, size=(4, 5, 6)).astype(np.int32)
    db = ph.NArray.from_numpy(b)
    for ax in range(3):
        other = tuple(i for i in range(3) if i != ax)
        assert np.array_equal(np.array(db.slices_sum(ax)), b.sum(axis=other).astype(np.float64))
    assert d.equals(d.clone()) and not d.equals(d + 1)


# ------------------------------------------------------------------ C4: heat
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_heat2d_bit_exact_small(ph, dt):
    rng = np.random.default_rng(3)
    for ny, nx in [(70, 300), (130, 257), (5, 1030)]:
        u0 = rng.random((ny, nx)).astype(dt)
        u = ph.NArray.from_numpy(u0)
        e = ph.heat2d(u, 0.2, 23, every=5)
        ref, eref = O.heat2d(u0, 0.2, 23, every=5)
        assert np.array_equal(u.to_numpy(), ref), (ny, nx)
        assert np.allclose(e, eref, rtol=1e-13, atol=0)


def test_heat2d_eigenmode_large(ph):
    ny, nx, p, q, r, steps = 1024, 2048, 1, 3, 0.2, 200
    i = np.arange(1, ny + 1)[:, None]
    j = np.arange(1, nx + 1)[None, :]
    u0 = np.sin(p * np.pi * i / (ny + 1)) * np.sin(q * np.pi * j / (nx + 1))
    lam = 1 - 4 * r * (np.sin(p * np.pi / (2 * (ny + 1))) ** 2 + np.sin(q * np.pi / (2 * (nx + 1))) ** 2)
    u = ph.NArray.from_numpy(u0)
    e = ph.heat2d(u, r, steps, every=100)
    assert np.max(np.abs(u.to_numpy() - lam ** steps * u0)) < 1e-12
    s0 = O.sum_exact(u0)
    for k, ek in enumerate(e):
        assert abs(ek - lam ** (100 * k) * s0) < 1e-10 * abs(s0)
