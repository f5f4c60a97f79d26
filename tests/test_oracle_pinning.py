"""Pins the CPU oracle before it is trusted as the checker of the device path:
against the known answers of ph-core's own specs (tests/golden/
reference_spec_vectors.json), against independent implementations (numpy
slicing / transposes / tile, Python's floored integer ops, math.fsum, numpy
matrix powers) and against closed forms (the FTCS eigenmode decay)."""
import json
import math
import os
import random

import numpy as np
import pytest

import oracle_api as O
from crystal import parse_literal, trace

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_spec_vectors.json")))


def P(lits):
    return [parse_literal(x) for x in lits]


# --------------------------------------------------------------------- ranges
def test_range_table_bound10():
    g = GOLD["ranges_valid_bound10"]
    for lit, f, s, l in g["cases"]:
        first, step, last, size = O.canonicalize_range(parse_literal(lit), g["bound"])
        assert (first, step, last) == (f, s, l), lit
        assert size == (l - f) // s + 1, lit  # spec/index_region_spec.cr:16


def test_range_out_of_bounds_and_step_conflict():
    for lit in GOLD["out_of_bounds_bound10"]["cases"]:
        with pytest.raises(O.OracleError) as e:
            O.region(P([lit]), [10])
        assert e.value.code == O.E_INDEX, lit
    for lit in GOLD["step_conflict"]["cases"]:
        with pytest.raises(O.OracleError) as e:
            O.region(P([lit]), [10])
        assert e.value.code == O.E_INDEX, lit
        with pytest.raises(O.OracleError) as e:
            O.region_absolute(P([lit]))
        assert e.value.code == O.E_INDEX, lit


def test_range_empty():
    for lit in GOLD["empty_bound10"]["cases"]:
        assert O.region(P([lit]), [10]).size == 0
        assert O.region_absolute(P([lit])).size == 0


def test_absolute_constructor():
    for lit, f, s, l in GOLD["fully_defined_bound10"]["cases"]:
        r = O.region_absolute(P([lit]))
        assert (r.first[0], r.step[0], r.last[0]) == (f, s, l), lit
    for lit in GOLD["implicit_bounds_absolute_raise"]["cases"]:
        with pytest.raises(O.OracleError):
            O.region_absolute(P([lit]))
    for lit in GOLD["negative_absolute_raise_index_error"]["cases"]:
        with pytest.raises(O.OracleError) as e:
            O.region_absolute(P([lit]))
        assert e.value.code == O.E_INDEX


def test_region_2d_includes_fits_translate_reverse():
    g = GOLD["region_2d"]
    r = O.region(P(g["literal"]), g["bound"])
    assert r.shape == g["shape"] and r.get("first") == g["first"]
    assert r.get("last") == g["last"] and r.get("step") == g["step"]
    with pytest.raises(O.OracleError):
        O.region(P(g["bad_literal"]), g["bound"])
    ra = O.region_absolute(P(g["absolute_literal"]))
    assert ra.shape == g["shape"] and ra.get("first") == g["first"] and ra.get("last") == g["last"]

    g = GOLD["includes"]
    r = O.region_absolute(P(g["literal"]))
    for c in g["outside"]:
        assert not O.region_includes(r, c), c
    for a in g["inside_a"]:
        for b in g["inside_b"]:
            assert O.region_includes(r, [a, b])

    g = GOLD["fits_in"]
    r = O.region_absolute(P(g["literal"]))
    assert O.region_fits_in(r, g["fits"]) and not O.region_fits_in(r, g["not_fits"])
    c = O.region_trim(O.region_cover(g["cover"]), g["trim_to"])
    assert O.region_fits_in(c, g["trim_to"]) and not O.region_fits_in(c, g["not_fits"])

    g = GOLD["translate"]
    r = O.region_absolute(P(g["literal"]))
    probe = O.local_to_absolute(r, g["probe"])
    t = O.region_translate(O.region_absolute(P(g["literal"])), g["offset"])
    assert t.shape == r.shape
    assert t.get("first") == [a + b for a, b in zip(r.get("first"), g["offset"])]
    assert O.local_to_absolute(t, g["probe"]) == [a + b for a, b in zip(probe, g["offset"])]
    for off in g["raise_offsets"]:
        with pytest.raises(O.OracleError) as e:
            O.region_translate(O.region_absolute(P(g["literal"])), off)
        assert e.value.code == O.E_INDEX

    g = GOLD["reverse"]
    r = O.region_absolute(P(g["literal"]))
    rv = O.region_reverse(O.region_absolute(P(g["literal"])))
    assert rv.get("first") == g["reversed_first"] == r.get("last")
    assert rv.get("last") == r.get("first")
    assert rv.get("step") == [-s for s in r.get("step")] and rv.shape == r.shape


def _random_literal(rng, bound):
    kind = rng.random()
    if kind < 0.15:
        return rng.randint(-bound - 1, bound)
    f = rng.choice([None, rng.randint(-bound - 1, bound + 1)])
    l = rng.choice([None, rng.randint(-bound - 1, bound + 1)])
    s = rng.choice([None, None, rng.choice([-3, -2, -1, 1, 2, 3, 4])])
    ex = rng.random() < 0.4
    fs = "" if f is None else str(f)
    ls = "" if l is None else str(l)
    dots = "..." if ex else ".."
    if s is None:
        return f"{fs}{dots}{ls}"
    return f"{fs}..{s}{dots}{ls}" if rng.random() < 0.7 else f"{fs}..({s}{dots}{ls})"


def test_range_properties_random():
    """Every canonical range is a valid arithmetic progression inside the axis,
    and its element set matches brute-force enumeration of the literal's
    integer set (Python range over the inferred endpoints)."""
    rng = random.Random(5)
    checked = 0
    for _ in range(3000):
        bound = rng.randint(0, 9)
        lit = _random_literal(rng, bound)
        try:
            f, s, l, n = O.canonicalize_range(parse_literal(lit), bound)
        except O.OracleError:
            continue
        checked += 1
        if n == 0:
            assert s == 0
            continue
        assert 0 <= f < bound and 0 <= l < bound
        assert l == f + s * (n - 1)
        # brute force: the integer set between the canonical endpoints
        p = parse_literal(lit)
        if p[0] == "index":
            continue
        _, pf, pl, ps, ex = p
        # implicit step: nil first is 0, nil last is bound-1, direction from
        # the endpoints; explicit step: nil endpoints follow the step's sign
        fc = 0 if pf is None else (bound + pf if pf < 0 else pf)
        if ps is None:
            lc = bound - 1 if pl is None else (bound + pl if pl < 0 else pl)
            step = 1 if lc >= fc else -1
        else:
            step = ps
            if pf is None:
                fc = 0 if step > 0 else bound - 1
            lc = (bound - 1 if step > 0 else 0) if pl is None else (bound + pl if pl < 0 else pl)
        stop = lc + (0 if (ex and pl is not None) else (1 if step > 0 else -1))
        assert list(range(fc, stop, step)) == [f + k * s for k in range(n)], lit
    assert checked > 500


# ------------------------------------------------------------ gather/scatter
def _np_select(r):
    """numpy index arrays for a region (proper dimensions)."""
    return np.ix_(*[[f + k * s for k in range(n)] for f, s, n in
                    zip(r.get("first"), r.get("step"), r.get("proper_shape"))])


def test_narray_spec_chunks():
    g = GOLD["narray_chunks"]
    stock = np.array(g["stock"], dtype=np.int32).reshape(g["stock_shape"])
    for lit, exp in g["fetch"]:
        r = O.region(P(lit), g["stock_shape"])
        assert O.gather(stock, r).reshape(-1).tolist() == exp
    assert O.region(P(g["fetch_empty"]), g["stock_shape"]).size == 0
    for lit, src, exp in g["set_src"]:
        d = stock.copy()
        O.scatter(d, O.region(P(lit), g["stock_shape"]), np.array(src, np.int32))
        assert d.reshape(-1).tolist() == exp
    for lit, v, exp in g["set_scalar"]:
        d = stock.copy()
        O.fill_region(d, O.region(P(lit), g["stock_shape"]), v)
        assert d.reshape(-1).tolist() == exp
    g = GOLD["multi_writable"]
    buf = np.array(g["buffer"], np.int32).reshape(g["shape"])
    d = buf.copy()
    O.fill_region(d, O.region(P(g["region"]), g["shape"]), g["scalar"])
    assert d.reshape(-1).tolist() == g["expected_scalar"]
    d = buf.copy()
    O.scatter(d, O.region(P(g["region"]), g["shape"]), np.array(g["chunk"], np.int32))
    assert d.reshape(-1).tolist() == g["expected_chunk"]


def test_gather_scatter_vs_numpy_random():
    rng = random.Random(11)
    for _ in range(300):
        rank = rng.randint(1, 4)
        shape = [rng.randint(1, 6) for _ in range(rank)]
        lits = [_random_literal(rng, b) for b in shape[: rng.randint(0, rank)]]
        try:
            r = O.region(P(lits), shape, drop=rng.random() < 0.7)
        except O.OracleError:
            continue
        src = np.arange(int(np.prod(shape)), dtype=np.int32).reshape(shape) * 7 + 3
        got = O.gather(src, r)
        if r.size == 0:
            assert got.size == 0
            continue
        exp = src[_np_select(r)].reshape(r.shape)
        assert np.array_equal(got, exp), (shape, lits)
        dst = np.zeros(shape, np.int32)
        vals = np.arange(r.size, dtype=np.int32) + 100
        O.scatter(dst, r, vals)
        ref = np.zeros(shape, np.int32)
        ref[_np_select(r)] = vals.reshape([int(v) for v in r.get("proper_shape")])
        assert np.array_equal(dst, ref), (shape, lits)


def test_transform_known_answers():
    g = GOLD["transforms"]
    assert O.chain_apply([("permute", g["permute"]["pattern"])], g["permute"]["coord"]) == g["permute"]["expected"]
    rs = g["reshape"]
    assert O.chain_apply([("reshape", rs["src_shape"], rs["new_shape"])], rs["coord"]) == rs["expected"]
    rv = g["reverse"]
    assert O.chain_apply([("reverse", rv["shape"])], rv["coord"]) == rv["expected"]


def _random_chain(rng, shape, max_len=4):
    """A random Region/Permute/Reverse/Reshape chain over `shape`, built
    alongside the same composition in numpy.  Returns (chain newest first,
    the view's shape, the view of np.arange(size) through numpy), so the last
    item holds, per view coordinate, the flat source index it maps to."""
    cur = np.arange(int(np.prod(shape)), dtype=np.int64).reshape(shape)
    chain = []  # newest first
    vshape = list(shape)
    for _ in range(rng.randint(1, max_len)):
        op = rng.choice(["region", "permute", "reverse", "reshape"])
        if op == "region":
            lits = [_random_literal(rng, b) for b in vshape]
            try:
                r = O.region(P(lits), vshape)
            except O.OracleError:
                continue
            if r.size == 0:
                continue
            cur = cur[_np_select(r)].reshape(r.shape)
            chain.insert(0, ("region", r))
            vshape = r.shape
        elif op == "permute":
            pat = list(range(len(vshape)))
            rng.shuffle(pat)
            cur = np.transpose(cur, pat)
            chain.insert(0, ("permute", pat))
            vshape = [vshape[p] for p in pat]
        elif op == "reverse":
            cur = cur[tuple(slice(None, None, -1) for _ in vshape)]
            chain.insert(0, ("reverse", list(vshape)))
        else:
            n = int(np.prod(vshape))
            divs = [d for d in range(1, n + 1) if n % d == 0]
            d = rng.choice(divs)
            ns = [d, n // d] if rng.random() < 0.5 else [n]
            cur = cur.reshape(ns)
            chain.insert(0, ("reshape", list(vshape), ns))
            vshape = ns
    return chain, list(vshape), cur


def test_view_chains_vs_numpy():
    """View#to_narr through region/permute/reverse/reshape chains equals the
    same composition done with numpy indexing."""
    rng = random.Random(3)
    for _ in range(200):
        rank = rng.randint(1, 3)
        shape = [rng.randint(1, 5) for _ in range(rank)]
        src = np.arange(int(np.prod(shape)), dtype=np.int64).reshape(shape)
        chain, vshape, cur = _random_chain(rng, shape)
        got = O.view_gather(src, chain, vshape)
        assert np.array_equal(got, cur), chain


def test_view_scatter_vs_numpy():
    """MutableView write-through (MultiWritable#unsafe_set_chunk over a view,
    src/multi_writable.cr:29-36, src/mutable_view.cr:16-18): the chain maps
    each view coordinate of the region to a flat destination index, found
    here by pushing np.arange(size) through the same composition in numpy;
    the source's elements go there in lexicographic order.  A chain applied
    oldest first, a colexicographic walk of the region or a region step
    dropped fails this."""
    rng = random.Random(11)
    checked = 0
    for _ in range(300):
        rank = rng.randint(1, 3)
        shape = [rng.randint(1, 6) for _ in range(rank)]
        chain, vshape, idx = _random_chain(rng, shape)
        lits = [_random_literal(rng, b) for b in vshape[: rng.randint(0, len(vshape))]]
        try:
            r = O.region(P(lits), vshape, drop=rng.random() < 0.5)
        except O.OracleError:
            continue
        dst = np.full(shape, -1, np.int32)
        if r.size == 0:
            assert np.array_equal(O.view_scatter(dst, chain, r, np.zeros(0, np.int32)), np.full(shape, -1))
            continue
        vals = (np.arange(r.size, dtype=np.int32) * 5 + 7)
        got = O.view_scatter(dst, chain, r, vals)
        targets = idx[_np_select(r)].reshape(-1)  # lexicographic over the region
        assert len(set(targets.tolist())) == targets.size  # a view chain is injective
        exp = np.full(int(np.prod(shape)), -1, np.int32)
        exp[targets] = vals
        assert np.array_equal(got.reshape(-1), exp), (shape, chain, lits)
        checked += 1
    assert checked > 150


def test_view_scatter_known_answers():
    """MultiWritable#set_chunk(region, chunk) (spec/multi_writable_spec.cr:
    24-33) written through MutableViews: the identity view with the spec's
    region, and a view restricted to that region then written whole.  Both
    must give the spec's buffer; the transposed spelling of the same write
    (a permuted view of the transposed array) must too."""
    g = GOLD["multi_writable"]
    shape, buf = g["shape"], g["buffer"]
    chunk = np.array(g["chunk"], np.int32).reshape(g["chunk_shape"])
    exp = g["expected_chunk"]
    r = O.region(P(g["region"]), shape)
    base = np.array(buf, np.int32).reshape(shape)
    assert O.view_scatter(base.copy(), [], r, chunk).reshape(-1).tolist() == exp
    whole = O.region_cover(r.shape)
    assert O.view_scatter(base.copy(), [("region", r)], whole, chunk).reshape(-1).tolist() == exp
    # the transposed array viewed through permute(1, 0) is the original layout
    t = np.ascontiguousarray(base.T)
    got = O.view_scatter(t, [("permute", [1, 0])], r, chunk)
    assert got.T.reshape(-1).tolist() == exp
    # and the scalar form (spec/multi_writable_spec.cr:14-22)
    full = np.full(r.size, g["scalar"], np.int32)
    assert O.view_scatter(base.copy(), [], r, full).reshape(-1).tolist() == g["expected_scalar"]


def test_tile_and_iteration_order():
    rng = np.random.default_rng(0)
    for shape, counts in [([1, 4], [3, 1]), ([2, 3], [2, 2]), ([3], [4]), ([2, 1, 2], [1, 3, 2])]:
        a = rng.integers(0, 100, size=shape).astype(np.int32)
        assert np.array_equal(O.tile(a, counts), np.tile(a, counts))
    g = GOLD["iteration_order"]
    a = np.array(g["values"], np.int32).reshape(g["shape"])
    r = O.region_cover(g["shape"])
    assert O.gather(a, r).reshape(-1).tolist() == g["lex"]
    colex = O.view_gather(a, [("permute", [1, 0])], [3, 3]).reshape(-1).tolist()
    assert colex == g["colex"]


# ------------------------------------------------------------------ masks
def test_set_mask_known_answers():
    """NArray#[]=(mask, value) and (mask, source) against the reference's
    spec (n_array_spec.cr:297-333), a mask of all false, and numpy.where on
    random masks of every element size."""
    m = GOLD["narray_masks"]
    mask = np.array(m["mask"]).reshape(2, 3)
    stock = np.arange(6, dtype=np.int32).reshape(2, 3)
    assert O.set_mask(stock.copy(), mask, m["scalar"]).reshape(-1).tolist() == m["expected_scalar"]
    assert O.set_mask(stock.copy(), mask, stock + m["src_plus"]).reshape(-1).tolist() == m["expected_src"]
    assert O.set_mask(stock.copy(), np.zeros((2, 3), bool), 99).tolist() == stock.tolist()
    rng = np.random.default_rng(5)
    for dt in (np.uint8, np.uint16, np.int32, np.float64):
        d = rng.integers(0, 100, (37, 41)).astype(dt)
        mk = rng.random((37, 41)) < 0.3
        src = rng.integers(100, 200, (37, 41)).astype(dt)
        assert np.array_equal(O.set_mask(d.copy(), mk, 7), np.where(mk, dt(7), d))
        assert np.array_equal(O.set_mask(d.copy(), mk, src), np.where(mk, src, d))


# ------------------------------------------------- map_with_index / coord
def test_map_with_index_and_coord_known_answers():
    """map_with_index is map_with over (element, Int32 lexicographic index)
    and map_with_coord over (element, coordinate per axis): the oracle's zip
    evaluator given np.arange / np.indices as the extra operands reproduces
    the reference specs (n_array_spec.cr:451-518) and the map_with_coord doc
    example (multi_indexable.cr:700-707)."""
    g = GOLD["narray_maps"]
    stock = np.arange(6, dtype=np.int32).reshape(2, 3)
    ci, cj = np.indices((2, 3), dtype=np.int32)
    r, fl = O.eval_expr(trace(lambda el, i, j: 10 * i + j + el, 3), [stock, ci, cj])
    assert r.reshape(-1).tolist() == g["with_coord"] and fl == 0
    idx = np.arange(6, dtype=np.int32).reshape(2, 3)
    r, _ = O.eval_expr(trace(lambda el, k: k + el, 2), [stock * 2, idx])
    assert r.reshape(-1).tolist() == g["times3"]
    doc = np.array(g["doc_coord_values"], np.int32).reshape(g["doc_coord_shape"])
    r, _ = O.eval_expr(trace(lambda el, i, j: el + (i + j), 3), [doc, ci, cj])
    assert r.reshape(-1).tolist() == g["doc_coord_sum"]
    # a transposed block order would give these instead
    bad, _ = O.eval_expr(trace(lambda el, i, j: 10 * i + j + el, 3), [stock, cj, ci])
    assert bad.reshape(-1).tolist() != g["with_coord"]
    # np.indices walks coordinates in the order ElemAndCoordIterator does
    # (elem_coord_iterator_spec.cr:10-22)
    o = GOLD["elem_coord_order"]
    v = np.array(o["values"], np.int32)
    pairs = [[int(e), [int(a), int(b)]] for e, a, b in zip(v.flat, *[c.flat for c in np.indices(v.shape)])]
    assert pairs == o["lex"]
    t = np.ascontiguousarray(v.T)
    pairs = [[int(e), [int(b), int(a)]] for e, a, b in zip(t.flat, *[c.flat for c in np.indices(t.shape)])]
    assert pairs == o["colex"]


# ---------------------------------------------------------------- arithmetic
def test_int32_known_answers_and_overflow():
    g = GOLD["narray_maps"]
    stock = np.arange(6, dtype=np.int32)
    r, fl = O.eval_expr(trace(lambda x: x ** 2), [stock])
    assert r.tolist() == g["squared"] and fl == 0
    r, fl = O.eval_expr(trace(lambda x: x * 3), [stock])
    assert r.tolist() == g["times3"]
    big = np.array([2 ** 31 - 1, -2 ** 31, 46341, 2, -7, 7], dtype=np.int32)
    rng = np.random.default_rng(1)
    for fn, pyop in [(lambda a, b: a + b, lambda a, b: a + b), (lambda a, b: a - b, lambda a, b: a - b),
                     (lambda a, b: a * b, lambda a, b: a * b)]:
        x = np.concatenate([big, rng.integers(-2 ** 31, 2 ** 31, 200).astype(np.int32)])
        y = np.concatenate([big[::-1], rng.integers(-2 ** 31, 2 ** 31, 200).astype(np.int32)])
        for i in range(x.size):
            res, fl = O.eval_expr(trace(fn, 2), [x[i:i + 1], y[i:i + 1]])
            exact = pyop(int(x[i]), int(y[i]))
            overflow = not (-2 ** 31 <= exact < 2 ** 31)
            assert bool(fl & 1) == overflow
            if not overflow:
                assert int(res[0]) == exact
    # floored // and % (Python's are floored), division by zero flagged
    x = rng.integers(-1000, 1000, 500).astype(np.int32)
    y = rng.integers(-30, 30, 500).astype(np.int32)
    y[y == 0] = 7
    q, fl = O.eval_expr(trace(lambda a, b: a // b, 2), [x, y])
    assert fl == 0 and q.tolist() == [int(a) // int(b) for a, b in zip(x, y)]
    m, _ = O.eval_expr(trace(lambda a, b: a % b, 2), [x, y])
    assert m.tolist() == [int(a) % int(b) for a, b in zip(x, y)]
    _, fl = O.eval_expr(trace(lambda a, b: a // b, 2), [np.array([1], np.int32), np.array([0], np.int32)])
    assert fl & 2
    # Int32 / Int32 -> Float64
    d, _ = O.eval_expr(trace(lambda a, b: a / b, 2), [x, y])
    assert d.dtype == np.float64 and np.array_equal(d, x.astype(np.float64) / y.astype(np.float64))
    # ** : negative exponent flagged; powers match Python's exact powers
    _, fl = O.eval_expr(trace(lambda a: a ** -1), [np.array([2], np.int32)])
    assert fl & 4
    bases = np.arange(-12, 13, dtype=np.int32)
    for e in range(0, 12):
        r, fl = O.eval_expr(trace(lambda a: a ** e), [bases])
        for b, v in zip(bases, r):
            exact = int(b) ** e
            if -2 ** 31 <= exact < 2 ** 31:
                assert int(v) == exact


def test_float_arithmetic_is_ieee():
    rng = np.random.default_rng(2)
    a64 = rng.standard_normal(1000)
    b64 = rng.standard_normal(1000)
    a32 = a64.astype(np.float32)
    b32 = b64.astype(np.float32)
    for fn in [lambda a, b: a + b, lambda a, b: a - b, lambda a, b: a * b, lambda a, b: a / b]:
        r, _ = O.eval_expr(trace(fn, 2), [a64, b64])
        assert np.array_equal(r, fn(a64, b64))
        r, _ = O.eval_expr(trace(fn, 2), [a32, b32])
        assert r.dtype == np.float32 and np.array_equal(r, fn(a32, b32))
        # S1: Float32 op Float64 -> Float64 (numpy promotes the same way)
        r, _ = O.eval_expr(trace(fn, 2), [a32, b64])
        assert r.dtype == np.float64 and np.array_equal(r, fn(a32.astype(np.float64), b64))
    # Float ** Int32 is binary powering: x**2 == x*x, x**3 == (x*x)*x
    r, _ = O.eval_expr(trace(lambda x: x ** 2), [a32])
    assert np.array_equal(r, a32 * a32)
    r, _ = O.eval_expr(trace(lambda x: x ** 3), [a64])
    assert np.array_equal(r, a64 * (a64 * a64)) or np.array_equal(r, (a64 * a64) * a64)
    # the ProcView block of the C3 workload, f32 input, S1 promotion at * 0.5
    r, _ = O.eval_expr(trace(lambda x: (x + 4) ** 2 * 0.5 - x), [a32])
    t = (a32 + np.float32(4)) * (a32 + np.float32(4))
    assert r.dtype == np.float64 and np.array_equal(r, t.astype(np.float64) * 0.5 - a32.astype(np.float64))


def test_reductions_vs_fsum():
    rng = np.random.default_rng(4)
    for n in [0, 1, 7, 1000, 100003]:
        x = rng.standard_normal(n) * 10 ** rng.uniform(-3, 3, n)
        assert O.sum_exact(x) == math.fsum(x.tolist())
        s = 0.0
        for v in x.tolist():
            s += v
        assert O.sum_seq(x) == s
        if n:
            assert O.max_seq(x) == max(x.tolist())
    # ill-conditioned: exact cancellation
    x = np.array([1e100, 1.0, -1e100, 1e-100], dtype=np.float64)
    assert O.sum_exact(x) == math.fsum(x.tolist()) == 1.0
    # non-finite terms, as math.fsum treats them (a NaN anywhere in a long
    # run returns at once instead of growing the partials)
    y = rng.standard_normal(200003)
    y[7] = np.nan
    assert math.isnan(O.sum_exact(y)) and math.isnan(math.fsum(y.tolist()))
    for vals in ([math.inf, 1.0, 2.0], [-math.inf, -math.inf, 5.0], [1.0, math.inf, math.inf]):
        assert O.sum_exact(np.array(vals)) == math.fsum(vals)
    assert math.isnan(O.sum_exact(np.array([math.inf, 3.0, -math.inf])))  # math.fsum raises here
    # Int32 checked sum: a prefix that overflows raises even if the total fits
    v, rc = O.sum_i32(np.array([2 ** 31 - 1, 1, -1], np.int32))
    assert rc == O.E_OVERFLOW and v == 2 ** 31 - 1
    v, rc = O.sum_i32(np.array([2 ** 31 - 1, -1, 1], np.int32))
    assert rc == 0 and v == 2 ** 31 - 1
    # slices(axis).map(&.sum): numpy per-axis sums of integer-valued data are exact
    a = rng.integers(-1000, 1000, size=(5, 6, 7)).astype(np.float64)
    for ax in range(3):
        other = tuple(i for i in range(3) if i != ax)
        assert np.array_equal(O.slices_reduce(a, ax, 0), a.sum(axis=other))
        assert np.array_equal(O.slices_reduce(a, ax, 2), a.sum(axis=other))
        assert np.array_equal(O.slices_reduce(a, ax, 1), a.max(axis=other))


# --------------------------------------------------------------------- heat
def test_heat1d_reference_example():
    """examples/heat_equation.cr as written: the update s' = M s is linear with
    a tridiagonal M; numpy's matrix power computes the same state."""
    n = 21
    coeff = (237 * 0.01) / (2700 * 900 * (0.05 * 0.05))
    M = np.eye(n)
    for i in range(1, n - 1):
        M[i, i - 1] += coeff
        M[i, i] -= 2 * coeff
        M[i, i + 1] += coeff
    M[0, 0] -= coeff
    M[0, 1] += coeff
    M[n - 1, n - 1] -= coeff
    M[n - 1, n - 2] += coeff
    s0 = np.full(n, 20.0)
    s0[0], s0[-1] = 0.0, 100.0
    for steps in [1, 2, 100, 10001]:
        got = O.heat1d_reference(steps)
        exp = np.linalg.matrix_power(M, steps) @ s0
        assert np.allclose(got, exp, rtol=1e-9, atol=1e-9), steps
        # the flux boundary updates conserve the total
        assert abs(got.sum() - s0.sum()) < 1e-9 * s0.sum()
        assert got.min() >= 0.0 and got.max() <= 100.0
    assert O.heat1d_reference().size == 21  # the full run: 10001 steps


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_heat2d_eigenmode(dt):
    """A discrete Dirichlet eigenmode decays by exactly lambda per step."""
    ny, nx, p, q, r, steps = 24, 40, 1, 3, 0.2, 30
    i = np.arange(1, ny + 1)[:, None]
    j = np.arange(1, nx + 1)[None, :]
    u0 = (np.sin(p * np.pi * i / (ny + 1)) * np.sin(q * np.pi * j / (nx + 1)))
    lam = 1 - 4 * r * (np.sin(p * np.pi / (2 * (ny + 1))) ** 2 + np.sin(q * np.pi / (2 * (nx + 1))) ** 2)
    u, e = O.heat2d(u0.astype(dt), r, steps, every=10)
    tol = 1e-12 if dt == np.float64 else 2e-5
    assert np.max(np.abs(u - lam ** steps * u0)) < tol * np.max(np.abs(u0))
    assert len(e) == 4
    for k, ek in enumerate(e):
        assert abs(ek - lam ** (10 * k) * u0.sum()) < (1e-11 if dt == np.float64 else 1e-5) * abs(u0.sum())
    # a wrong sign on one neighbour breaks this: the 4-neighbour sum is needed
    assert lam < 1


def test_heat2d_symmetry_and_energy_decay():
    rng = np.random.default_rng(7)
    a = rng.random((16, 16))
    u0 = a + a[::-1, :] + a[:, ::-1] + a[::-1, ::-1]
    u, e = O.heat2d(u0, 0.2, 40, every=5)
    # mirror symmetry holds up to rounding (the neighbour sum is ordered N, S, W, E)
    assert np.allclose(u, u[::-1, :], rtol=1e-13, atol=0) and np.allclose(u, u[:, ::-1], rtol=1e-13, atol=0)
    assert all(e[k + 1] < e[k] for k in range(len(e) - 1))  # zero boundary drains heat


# ---------------------------------------------------------------- generator
def _splitmix(seed, k):
    M = (1 << 64) - 1
    z = (seed + (k + 1) * 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_generator_is_splitmix64():
    for seed in [0, 1, 2, 3, 4, 12345]:
        for k in [0, 1, 2, 1000, 2 ** 40]:
            assert O.draw(seed, k) == _splitmix(seed, k)
    # the standard splitmix64 stream from state 0: first output
    assert _splitmix(0, 0) == 0xE220A8397B1DCDAF
    u = O.fill_random([5], O.DT_F64, 0, 0, 0, -1.0, 1.0)
    assert u.tolist() == [-1.0 + 2.0 * ((_splitmix(0, k) >> 11) * 2.0 ** -53) for k in range(5)]
    ints = O.fill_random([5], O.DT_I32, 1, 0, 2)
    assert ints.tolist() == [_splitmix(1, k) >> 33 for k in range(5)]
    nrm = O.fill_random([4], O.DT_F64, 2, 0, 1, 0.0, 1.0)
    for k in range(4):
        kp = k & ~1
        u1 = (_splitmix(2, kp) >> 11) * 2.0 ** -53
        u2 = (_splitmix(2, kp + 1) >> 11) * 2.0 ** -53
        r = math.sqrt(-2.0 * math.log(1.0 - u1))
        t = 2 * math.pi * u2
        assert nrm[k] == pytest.approx(r * (math.sin(t) if k & 1 else math.cos(t)), rel=1e-14)


def test_concatenate_vs_numpy():
    """or_concatenate (concatenate_to_slice's chunk loop) against
    numpy.concatenate on random shapes, axes and element sizes."""
    rng = np.random.default_rng(12)
    for _ in range(200):
        rank = int(rng.integers(1, 5))
        base = [int(v) for v in rng.integers(0, 5, rank)]
        axis = int(rng.integers(0, rank))
        dt = [np.int32, np.float64, np.uint16, np.uint8][int(rng.integers(0, 4))]
        arrs = []
        for _ in range(int(rng.integers(1, 5))):
            s = list(base)
            s[axis] = int(rng.integers(0, 5))
            arrs.append(rng.integers(0, 200, s).astype(dt))
        assert np.array_equal(O.concatenate(arrs, axis), np.concatenate(arrs, axis=axis))


def test_max_follows_enumerable_greater_than():
    """Enumerable#max keeps the first element and replaces it only by a
    greater one (`>`, src/multi_indexable.cr:18 includes Enumerable): a
    leading NaN is kept (nothing compares greater than it), a later NaN is
    skipped, and of equal zeros the first one's sign survives.  Known
    answers from that definition."""
    nan = float("nan")
    assert math.isnan(O.max_seq(np.array([nan, 1.0, 5.0])))
    assert O.max_seq(np.array([1.0, nan, 2.0])) == 2.0
    assert O.max_seq(np.array([3.0, nan])) == 3.0
    assert math.copysign(1, O.max_seq(np.array([-0.0, 0.0]))) == -1
    assert math.copysign(1, O.max_seq(np.array([0.0, -0.0]))) == 1
    assert math.copysign(1, O.max_seq(np.array([-1.0, -0.0, 0.0, -2.0]))) == -1
    assert math.copysign(1, O.max_seq(np.array([-3.0, 0.0, -0.0]))) == 1
    s = O.slices_reduce(np.array([[nan, 2.0], [-0.0, 0.0]]), 0, 1)
    assert math.isnan(s[0]) and math.copysign(1, s[1]) == -1
    s = O.slices_reduce(np.array([[nan, 2.0], [-0.0, 0.0]]), 1, 1)
    assert math.isnan(s[0]) and s[1] == 2.0


def test_sequential_sum_in_pieces():
    """A sequential sum fed in pieces equals the one-shot sequential sum (the
    same additions in the same order), and differs from a pairwise sum on
    data where the order shows (absorption of small terms)."""
    rng = np.random.default_rng(12)
    x = np.concatenate([[1e16], rng.random(1000), [-1e16], rng.random(1000)])
    s = 0.0
    for k in range(0, x.size, 137):
        s = O.sum_seq_from(x[k:k + 137], s)
    assert s == O.sum_seq(x)
    acc = 0.0
    for v in x:
        acc += v
    assert s == acc
    assert s != math.fsum(x)
