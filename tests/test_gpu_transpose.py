"""Permuted views at sizes that take the LDS-tiled transposing copy (the
source contiguous along one axis, the destination along another): gathers
of permuted, reversed and region views, and scatters through permuted
MutableViews, for 1-, 2-, 4- and 8-byte elements, against the oracle's view
gather and scatter (src/view_util/transforms.cr:118-305)."""
import numpy as np
import pytest

import oracle_api as O
from crystal import parse_literal

pytestmark = pytest.mark.gpu

CASES = [
    ((64, 64), [1, 0]), ((100, 37), [1, 0]), ((16, 2000), [1, 0]), ((17, 16), [1, 0]),
    ((3, 70, 130), [2, 0, 1]), ((3, 70, 130), [1, 2, 0]), ((70, 3, 130), [2, 1, 0]),
    ((5, 40, 33), [0, 2, 1]), ((2, 1, 300, 17), [3, 2, 1, 0]), ((1000, 1000), [1, 0]),
]
KINDS = ["i32", "f64", "bf16", "bool", "f32"]


def _host(rng, shape, kind):
    if kind == "i32":
        return rng.integers(-2 ** 31, 2 ** 31, shape, dtype=np.int64).astype(np.int32)
    if kind == "f64":
        return rng.standard_normal(shape)
    if kind == "f32":
        return rng.standard_normal(shape).astype(np.float32)
    if kind == "bf16":
        return rng.integers(0, 2 ** 16, shape, dtype=np.int64).astype(np.uint16)
    return rng.random(shape) < 0.5


def _dev(ph, host, kind):
    if kind == "bf16":
        return ph.NArray.from_numpy(host, ph.DType.BF16)
    return ph.NArray.from_numpy(host)


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint8) if a.dtype == np.bool_ else a.view(f"u{a.itemsize}")


@pytest.mark.parametrize("kind", KINDS)
def test_permuted_gathers(ph, kind):
    rng = np.random.default_rng(51)
    for shape, perm in CASES:
        host = _host(rng, shape, kind)
        dev = _dev(ph, host, kind)
        vshape = [shape[p] for p in perm]
        got = dev.permute(*perm).to_narr().to_numpy()
        exp = O.view_gather(host, [("permute", perm)], vshape)
        assert np.array_equal(_bits(got), _bits(exp)), (shape, perm, kind)
        # reversed, then permuted: descending unit strides on both sides
        chain = [("permute", perm), ("reverse", list(shape))]
        got = dev.view().reverse().permute(perm).to_narr().to_numpy()
        assert np.array_equal(_bits(got), _bits(O.view_gather(host, chain, vshape))), (shape, perm, kind)


def test_region_then_permute(ph):
    rng = np.random.default_rng(53)
    host = rng.integers(-1000, 1000, (300, 257), dtype=np.int64).astype(np.int32)
    dev = ph.NArray.from_numpy(host)
    for lits in [("3..290", ".."), ("..", "5..200"), ("-1..0", "1..2..256"), ("10..3..299", "-2..-1..0")]:
        r = O.region([parse_literal(x) for x in lits], list(host.shape))
        chain = [("permute", [1, 0]), ("region", r)]
        got = dev.view(*lits).permute(1, 0).to_narr().to_numpy()
        exp = O.view_gather(host, chain, [r.shape[1], r.shape[0]])
        assert np.array_equal(got, exp), lits


@pytest.mark.parametrize("kind", ["i32", "f64", "bf16", "bool"])
def test_scatter_through_permuted_mutable_view(ph, kind):
    rng = np.random.default_rng(57)
    for shape, perm in CASES[:9]:
        host = _host(rng, shape, kind)
        vshape = [shape[p] for p in perm]
        src = _host(rng, vshape, kind)
        target = _dev(ph, host, kind)
        mv = target.mutable_view().permute(*perm)
        mv[[]] = _dev(ph, src, kind)
        ref = O.view_scatter(host.copy(), [("permute", perm)], O.region_cover(vshape), src)
        assert np.array_equal(_bits(target.to_numpy()), _bits(ref)), (shape, perm, kind)


def test_large_transpose(ph):
    """8192 x 8192 Int32, whole and offset by a region: every element checked
    against numpy's transpose (which the oracle's view gather is pinned to)."""
    x = ph.NArray.random([8192, 8192], ph.DType.I32, seed=1, dist="int31")
    h = x.to_numpy()
    assert np.array_equal(x.permute(1, 0).to_narr().to_numpy(), h.T)
    assert np.array_equal(x.view("1..8190", "7..").permute(1, 0).to_narr().to_numpy(), h[1:8191, 7:].T)


def test_maps_over_transposed_operands(ph):
    """Maps whose operands are transposed views: one or two are staged
    through LDS inside the map itself (ph_map_tiled); with three, two are
    copied through the tiled transpose first and the third is read in place;
    bit-exact against the oracle either way."""
    from crystal import trace
    rng = np.random.default_rng(59)
    hs = [rng.integers(-2 ** 31, 2 ** 31, (300, 257), dtype=np.int64).astype(np.int32) for _ in range(3)]
    c = rng.integers(-2 ** 31, 2 ** 31, (257, 300), dtype=np.int64).astype(np.int32)
    xs = [ph.NArray.from_numpy(h).view().permute(1, 0) for h in hs]
    ts = [np.ascontiguousarray(h.T) for h in hs]
    dc = ph.NArray.from_numpy(c)
    fn2 = lambda a, b: a.wadd(b)  # noqa: E731
    exp, _ = O.eval_expr(trace(fn2, 2), [ts[0], c])
    assert np.array_equal(xs[0].map_with(dc, block=fn2).to_numpy(), exp)
    assert np.array_equal(dc.map_with(xs[0], block=lambda a, b: b.wadd(a)).to_numpy(), exp)
    fn4 = lambda a, b, d, e: a.wadd(b).wsub(d) ^ e  # noqa: E731
    exp, _ = O.eval_expr(trace(fn4, 4), [ts[0], ts[1], ts[2], c])
    assert np.array_equal(xs[0].map_with(xs[1], xs[2], dc, block=fn4).to_numpy(), exp)
    # a transposed float view through a ProcView block
    f = rng.standard_normal((300, 257)).astype(np.float32)
    pv = ph.NArray.from_numpy(f).view().permute(1, 0).process(lambda v: (v + 4) ** 2 * 0.5 - v)
    exp, _ = O.eval_expr(trace(lambda v: (v + 4) ** 2 * 0.5 - v), [np.ascontiguousarray(f.T)])
    assert np.array_equal(pv.to_narr().to_numpy(), exp)


def _one_launch(ph, make):
    """The materialised result of make() and the kernel launches it took."""
    r = make()
    n0 = ph.launch_count()
    r = r.materialize()
    return r, ph.launch_count() - n0


def test_tiled_map_layouts(ph):
    """The LDS-tiled map (one launch, no staging copy) beyond one transposed
    Int32 operand: two transposed Float64 operands (two 64 x 65 LDS tiles,
    66.5 KB), a rank-3 permuted view (the loop over the outer axis), a
    transposed view that is also reversed (descending unit strides), and
    extents that are not multiples of 64; then an Int32 division by a zero
    read through the tiled path raises DivisionByZeroError.  Bit-exact
    against the oracle."""
    from crystal import trace
    rng = np.random.default_rng(61)
    # two transposed Float64 operands and a contiguous one
    hs = [rng.standard_normal((300, 257)) for _ in range(2)]
    c = rng.standard_normal((257, 300))
    xs = [ph.NArray.from_numpy(h).view().permute(1, 0) for h in hs]
    dc = ph.NArray.from_numpy(c)
    fn = lambda a, b, d: (a + b) * d - a  # noqa: E731
    exp, _ = O.eval_expr(trace(fn, 3), [np.ascontiguousarray(hs[0].T), np.ascontiguousarray(hs[1].T), c])
    got, n = _one_launch(ph, lambda: xs[0].map_with(xs[1], dc, block=fn))
    assert n == 1, n
    assert np.array_equal(got.to_numpy().view(np.uint64), exp.view(np.uint64))
    # rank 3: (5, 40, 70) viewed as (5, 70, 40); the outer axis is looped
    h3 = rng.integers(-2 ** 31, 2 ** 31, (5, 40, 70), dtype=np.int64).astype(np.int32)
    c3 = rng.integers(-2 ** 31, 2 ** 31, (5, 70, 40), dtype=np.int64).astype(np.int32)
    fn2 = lambda a, b: a.wadd(b) ^ (a & 4095)  # noqa: E731
    exp, _ = O.eval_expr(trace(fn2, 2), [np.ascontiguousarray(h3.transpose(0, 2, 1)), c3])
    got, n = _one_launch(ph, lambda: ph.NArray.from_numpy(h3).view().permute(0, 2, 1).map_with(
        ph.NArray.from_numpy(c3), block=fn2))
    assert n == 1, n
    assert np.array_equal(got.to_numpy(), exp)
    # transposed and reversed on both axes, 200 x 131
    hr = rng.integers(-2 ** 31, 2 ** 31, (200, 131), dtype=np.int64).astype(np.int32)
    cr = rng.integers(-2 ** 31, 2 ** 31, (131, 200), dtype=np.int64).astype(np.int32)
    vr = np.ascontiguousarray(hr[::-1, ::-1].T)
    exp, _ = O.eval_expr(trace(fn2, 2), [vr, cr])
    got, n = _one_launch(ph, lambda: ph.NArray.from_numpy(hr).view().reverse().permute(1, 0).map_with(
        ph.NArray.from_numpy(cr), block=fn2))
    assert n == 1, n
    assert np.array_equal(got.to_numpy(), exp)
    # an Int32 divisor with a zero, read through the tiled path
    num = rng.integers(-1000, 1000, (131, 200), dtype=np.int64).astype(np.int32)
    den = rng.integers(1, 50, (200, 131), dtype=np.int64).astype(np.int32)
    xden = ph.NArray.from_numpy(den).view().permute(1, 0)
    exp, fl = O.eval_expr(trace(lambda a, b: a // b, 2), [num, np.ascontiguousarray(den.T)])
    assert fl == 0
    assert np.array_equal(ph.NArray.from_numpy(num).map_with(xden, block=lambda a, b: a // b).to_numpy(), exp)
    den[77, 13] = 0
    xden = ph.NArray.from_numpy(den).view().permute(1, 0)
    with pytest.raises(ph.DivisionByZeroError):
        ph.NArray.from_numpy(num).map_with(xden, block=lambda a, b: a // b).to_numpy()


def test_reductions_over_transposed_views_int32(ph):
    """Int32 sum / max in full and along each axis, and the six-result
    pass, over a transposed view: the view is copied through the tiled
    transpose first (the overflow check follows its lexicographic order),
    so the results equal those of its materialised copy bit for bit, and
    numpy's exact sums."""
    rng = np.random.default_rng(61)
    h = rng.integers(-1000, 1000, (300, 1000)).astype(np.int32)
    v = ph.NArray.from_numpy(h).view().permute(1, 0)
    m = v.to_narr()
    assert v.sum() == m.sum() and v.max() == m.max()
    for ax in (0, 1):
        assert list(v.slices_sum(ax)) == list(m.slices_sum(ax))
        assert list(v.slices_max(ax)) == list(m.slices_max(ax))
    assert v.reduce2d_all() == m.reduce2d_all()
    assert v.sum() == int(h.astype(np.int64).sum())
    assert np.array_equal(np.array(v.slices_sum(0)), h.T.astype(np.int64).sum(axis=1))
    # the view's lexicographic running total leaves Int32 (MAX, then +5),
    # the storage's order would not (MAX, -10, ..., +5)
    big = np.zeros((40, 30), np.int32)
    big[0, 0] = 2 ** 31 - 1
    big[0, 1] = 5
    big[1, 0] = -10
    with pytest.raises(ph.OverflowError):
        ph.NArray.from_numpy(np.ascontiguousarray(big.T)).view().permute(1, 0).sum()


def test_reductions_over_transposed_views_float(ph):
    """Float sums over a transposed view run in the storage's orientation
    (rows and columns trade places): maxima equal the materialised copy's
    bit for bit, sums agree with the oracle's exact sums to 1e-12."""
    rng = np.random.default_rng(63)
    h = rng.standard_normal((300, 1000)).astype(np.float32)
    v = ph.NArray.from_numpy(h).view().permute(1, 0)
    vt = np.ascontiguousarray(h.T).astype(np.float64)
    m = v.to_narr()
    assert v.max() == m.max()
    assert abs(v.sum() - O.sum_exact(vt)) <= 1e-12 * np.abs(vt).sum()
    for ax in (0, 1):
        assert list(v.slices_max(ax)) == list(m.slices_max(ax))
        want = O.slices_reduce(vt if ax == 0 else np.ascontiguousarray(vt.T), 0, 2)
        scale = np.abs(vt).sum(axis=1 - ax)
        assert np.all(np.abs(np.array(v.slices_sum(ax)) - want) <= 1e-12 * scale), ax
    r, rm = v.reduce2d_all(), m.reduce2d_all()
    for k in ("full_max", "row_max", "col_max"):
        assert r[k] == rm[k], k
    for k in ("row_sum", "col_sum"):
        assert np.allclose(r[k], rm[k], rtol=1e-12, atol=0), k
    assert abs(r["full_sum"] - rm["full_sum"]) <= 1e-12 * np.abs(vt).sum()


def test_middle_axis_reduction_of_reshaped_view(ph):
    """slices(k) for a middle axis k of a view with a reshape stage."""
    rng = np.random.default_rng(67)
    h = rng.integers(-1000, 1000, (4, 6, 10)).astype(np.int32)
    x = ph.NArray.from_numpy(h)
    # (4, 3, 5) with strides (60, 20, 2): its last two axes do not merge, so
    # the reshape to (4, 5, 3) needs a stage
    v = x.view("..", "..2..", "..2..").reshape(4, 5, 3)
    hv = h[:, ::2, ::2].reshape(4, 5, 3)
    for ax in range(3):
        others = tuple(i for i in range(3) if i != ax)
        assert np.array_equal(np.array(v.slices_sum(ax)), hv.astype(np.int64).sum(axis=others)), ax
        assert np.array_equal(np.array(v.slices_max(ax)), hv.max(axis=others)), ax


def test_tiled_map_more_layouts(ph):
    """The LDS-tiled map over extents that are multiples of 4 but not of 64
    (partial tiles at the edges), three inputs of mixed element types with
    Float32 specials, a rank-3 view, regions with aligned and unaligned
    offsets, and an Int32 division by a zero that raises.  One launch each,
    bit-exact against the oracle."""
    from crystal import trace
    rng = np.random.default_rng(62)
    i32 = lambda s: rng.integers(-2 ** 31, 2 ** 31, s, dtype=np.int64).astype(np.int32)  # noqa: E731
    h, c = i32((300, 256)), i32((256, 300))
    fn2 = lambda a, b: a.wadd(b) ^ (a & 4095)  # noqa: E731
    exp, _ = O.eval_expr(trace(fn2, 2), [np.ascontiguousarray(h.T), c])
    got, n = _one_launch(ph, lambda: ph.NArray.from_numpy(h).view().permute(1, 0).map_with(
        ph.NArray.from_numpy(c), block=fn2))
    assert n == 1 and np.array_equal(got.to_numpy(), exp)
    # Float32 transposed, Int32 and Float32 contiguous: 200 x 132
    f = rng.standard_normal((132, 200)).astype(np.float32)
    f[3, 5], f[7, 9], f[100, 0] = np.inf, -0.0, np.nan
    gi = rng.integers(-50, 50, (200, 132)).astype(np.int32)
    gf = rng.standard_normal((200, 132)).astype(np.float32)
    fn3 = lambda a, b, d: (a * b - d) / 2  # noqa: E731 (Float32 throughout)
    exp, _ = O.eval_expr(trace(fn3, 3), [np.ascontiguousarray(f.T), gi, gf])
    got, n = _one_launch(ph, lambda: ph.NArray.from_numpy(f).view().permute(1, 0).map_with(
        ph.NArray.from_numpy(gi), ph.NArray.from_numpy(gf), block=fn3))
    assert n == 1 and exp.dtype == np.float32 and got.to_numpy().dtype == exp.dtype
    assert np.array_equal(got.to_numpy().view(np.uint32), exp.view(np.uint32))
    # rank 3: (3, 64, 72) viewed as (3, 72, 64)
    h3, c3 = i32((3, 64, 72)), i32((3, 72, 64))
    exp, _ = O.eval_expr(trace(fn2, 2), [np.ascontiguousarray(h3.transpose(0, 2, 1)), c3])
    got, n = _one_launch(ph, lambda: ph.NArray.from_numpy(h3).view().permute(0, 2, 1).map_with(
        ph.NArray.from_numpy(c3), block=fn2))
    assert n == 1 and np.array_equal(got.to_numpy(), exp)
    # regions: rows 4.. (offset 4 x 256, aligned) and columns 1.. (offset 1)
    for lits, sub in [(("4..", ".."), h[4:, :]), (("..", "1..-4"), h[:, 1:-3])]:
        cs = i32((sub.shape[1], sub.shape[0]))
        exp, _ = O.eval_expr(trace(fn2, 2), [np.ascontiguousarray(sub.T), cs])
        got, n = _one_launch(ph, lambda: ph.NArray.from_numpy(h).view(*lits).permute(1, 0).map_with(
            ph.NArray.from_numpy(cs), block=fn2))
        assert n == 1 and np.array_equal(got.to_numpy(), exp), lits
    # Int32 // with a zero read through the transposed operand
    num = rng.integers(-1000, 1000, (128, 200)).astype(np.int32)
    den = rng.integers(1, 50, (200, 128)).astype(np.int32)
    exp, fl = O.eval_expr(trace(lambda a, b: a // b, 2), [num, np.ascontiguousarray(den.T)])
    assert fl == 0
    xden = ph.NArray.from_numpy(den).view().permute(1, 0)
    assert np.array_equal(ph.NArray.from_numpy(num).map_with(xden, block=lambda a, b: a // b).to_numpy(), exp)
    den[77, 13] = 0
    xden = ph.NArray.from_numpy(den).view().permute(1, 0)
    with pytest.raises(ph.DivisionByZeroError):
        ph.NArray.from_numpy(num).map_with(xden, block=lambda a, b: a // b).to_numpy()
