"""Edge cases of the device path against the oracle: tiny and ragged heat
grids (every steps-per-pass remainder), reductions over ragged shapes of
every element type and axis, Int32 running-sum overflow, empty inputs, and
arrays beyond 2^31 elements (64-bit indexing)."""
import numpy as np
import pytest

import oracle_api as O
from crystal import parse_literal


def O_lit(lit):
    """(first, last, step, exclusive) of a range literal."""
    return parse_literal(lit)[1:]

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_heat_tiny_and_ragged_grids(ph, dt):
    rng = np.random.default_rng(17)
    for ny, nx in [(1, 1), (1, 7), (7, 1), (2, 3), (3, 1030), (130, 5), (65, 257), (200, 249)]:
        for steps in (1, 2, 3, 9, 17):
            u0 = rng.random((ny, nx)).astype(dt)
            u = ph.NArray.from_numpy(u0)
            e = ph.heat2d(u, 0.2, steps, every=4)
            ref, eref = O.heat2d(u0, 0.2, steps, every=4)
            assert np.array_equal(u.to_numpy(), ref), (ny, nx, steps)
            assert np.allclose(e, eref, rtol=1e-13, atol=1e-300), (ny, nx, steps)


def _device(ph, a, storage):
    """The device array and the values it holds (BF16: the truncated bits)."""
    if storage == "BF16":
        bits = (a.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
        return ph.NArray.from_numpy(bits, ph.DType.BF16), (bits.astype(np.uint32) << 16).view(np.float32)
    return ph.NArray.from_numpy(a), a


@pytest.mark.parametrize("storage", ["F64", "F32", "BF16", "I32"])
def test_reductions_ragged_shapes(ph, storage):
    rng = np.random.default_rng(23)
    np_dt = {"F64": np.float64, "F32": np.float32, "BF16": np.float32, "I32": np.int32}[storage]
    for shape in [(1,), (5,), (1, 1), (1, 1027), (1027, 1), (513, 1029), (3, 5, 7), (2, 1, 1030), (17, 3, 2, 9),
                  (100, 1000), (37, 2056), (3, 8), (1100, 3000)]:
        if storage == "I32":
            a = rng.integers(-1000, 1000, shape).astype(np.int32)
        else:
            a = rng.standard_normal(shape).astype(np_dt)
        d, host = _device(ph, a, storage)
        h64 = host.astype(np.float64)
        # full sum and max
        s = d.sum()
        if storage == "I32":
            assert s == int(a.astype(np.int64).sum())
        else:
            ex = O.sum_exact(h64)
            assert abs(s - ex) <= 1e-12 * np.abs(h64).sum() + 1e-30, (shape, s, ex)
        assert float(d.max()) == float(h64.max()), shape
        # slices(k).map(&.sum) / (&.max) along every axis
        for ax in range(len(shape)):
            others = tuple(i for i in range(len(shape)) if i != ax)
            ss = np.array(d.slices_sum(ax))
            sm = np.array(d.slices_max(ax))
            assert np.array_equal(sm, h64.max(axis=others) if others else h64), (shape, ax)
            if storage == "I32":
                assert np.array_equal(ss, h64.sum(axis=others) if others else h64), (shape, ax)
            else:
                want = np.array([O.sum_exact(np.take(h64, j, axis=ax)) for j in range(shape[ax])])
                scale = np.array([np.abs(np.take(h64, j, axis=ax)).sum() for j in range(shape[ax])])
                assert np.all(np.abs(ss - want) <= 1e-12 * scale + 1e-30), (shape, ax)


def test_int32_running_sum_overflow(ph):
    """Enumerable#sum with checked Int32 `+` raises as soon as the running sum
    leaves the Int32 range, even if later terms bring it back."""
    big = 2 ** 31 - 1
    ok = ph.NArray.from_numpy(np.array([big - 5, 5, -7], np.int32))
    assert ok.sum() == big - 7
    back = ph.NArray.from_numpy(np.array([big - 5, 6, -7], np.int32))  # crosses, then returns
    with pytest.raises(ph.OverflowError):
        back.sum()
    low = ph.NArray.from_numpy(np.array([-2 ** 31, 1, -1, -1], np.int32))
    with pytest.raises(ph.OverflowError):
        low.sum()
    # across a tile boundary of a 2-D reduction
    a = np.zeros((40, 3000), np.int32)
    a[0, 0] = big
    a[20, 2999] = 1
    a[39, 0] = -5
    with pytest.raises(ph.OverflowError):
        ph.NArray.from_numpy(a).sum()
    a[20, 2999] = -1
    assert ph.NArray.from_numpy(a).sum() == big - 6


def test_empty_inputs(ph):
    e = ph.NArray.fill([0, 5], 1.5)
    assert (e + 1.0).to_numpy().shape == (0, 5)
    assert e.sum() == 0.0
    with pytest.raises(ph.EmptyError):
        e.max()
    assert list(e.slices_sum(1)) == [0.0] * 5
    # any range over an empty axis raises, as canonicalize_range does
    # (range_syntax.cr:145-150); the whole-array view is empty
    for lit in ["..", "0...0"]:
        with pytest.raises(O.OracleError):
            O.canonicalize_range(("range",) + tuple(O_lit(lit)), 0)
        with pytest.raises(IndexError):
            e.view(lit, "1..2")
    assert e.view().to_narr().size == 0 and e.view().to_narr().shape == [0, 5]
    z = ph.NArray.fill([0], 0)
    assert z.sum() == 0
    h = ph.NArray.fill([4, 4], 1.0)
    assert np.array_equal(np.array(ph.heat2d(h, 0.2, 0, every=1)), [16.0])  # zero steps: energy only


def test_beyond_2_31_elements(ph):
    """A 2^31 + 4100-element Int32 array (8.6 GB): strided gather and scatter
    with offsets past 2^31 take the 64-bit index path."""
    n = 2 ** 31 + 4100
    x = ph.NArray.random([n], ph.DType.I32, seed=1, dist="int31")
    xh = x.to_numpy()
    g = x.view("2147480000..3..").to_narr().to_numpy()
    assert np.array_equal(g, xh[2147480000::3])
    tail = ph.NArray.from_numpy(np.arange(1000, dtype=np.int32))
    x[["-2000..2.."]] = tail
    got = x.to_numpy()
    exp = xh.copy()
    exp[n - 2000::2] = np.arange(1000, dtype=np.int32)
    assert np.array_equal(got[n - 4096:], exp[n - 4096:]) and np.array_equal(got[:4096], exp[:4096])
    assert x.view("-3..").to_narr().to_numpy().tolist() == exp[-3:].tolist()


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_heat_edge_placed_sweeps(ph, dt):
    """Grids whose rows allow whole-vector lanes: strips flush with the west
    and east edges, the top band swept down and the bottom band up from the
    grid's edge, two and eight steps per pass (17 = 8 + 8 + 1), bit-exact."""
    rng = np.random.default_rng(29)
    for ny, nx in [(40, 256), (300, 1000), (1000, 2048), (97, 260)]:
        for steps in (2, 8, 17):
            u0 = rng.random((ny, nx)).astype(dt)
            u = ph.NArray.from_numpy(u0)
            e = ph.heat2d(u, 0.2, steps, every=steps)
            ref, eref = O.heat2d(u0, 0.2, steps, every=steps)
            assert np.array_equal(u.to_numpy(), ref), (ny, nx, steps)
            assert np.allclose(e, eref, rtol=1e-13, atol=1e-300), (ny, nx, steps)


def test_exact_product_fusion_edge_values(ph):
    """Float32 or Int32 values widened to Float64, times a power of two, then
    added or subtracted: the device fuses these into one multiply-add, which
    must equal the two roundings of the oracle at subnormals, the extremes,
    zeros of both signs, infinities and NaN."""
    from crystal import trace
    f = np.float32
    edge = np.array([0.0, -0.0, 1.4e-45, -1.4e-45, 1.1754942e-38, 1.17549435e-38, -1.17549435e-38,
                     3.4028235e38, -3.4028235e38, np.inf, -np.inf, np.nan, 1.0, -2.5, 1e-30, 7e37], np.float32)
    rng = np.random.default_rng(31)
    x = np.concatenate([edge, rng.standard_normal(1000).astype(f), (rng.standard_normal(300) * 1e-40).astype(f)])
    w = np.concatenate([edge[::-1], rng.standard_normal(1000 + 300).astype(f)])
    assert x.size == w.size
    cases = [
        (lambda v: v * 0.5 - 1.0, 1),
        (lambda v, u: u - v * 0.25, 2),
        (lambda v, u: v * 1024.0 + u, 2),
        (lambda v: v * (2.0 ** -800) + 1e-300, 1),
        (lambda v: v * (2.0 ** 890) - 1e300, 1),
        (lambda v: (v + 4) ** 2 * 0.5 - v, 1),
    ]
    for fn, nin in cases:
        ins = [x, w][:nin]
        exp, _ = O.eval_expr(trace(fn, nin), ins)
        dev = [ph.NArray.from_numpy(a) for a in ins]
        got = (dev[0].map_with(*dev[1:], block=fn) if nin > 1 else dev[0].process(fn).to_narr()).to_numpy()
        assert got.dtype == exp.dtype == np.float64
        assert np.array_equal(got, exp, equal_nan=True)
        assert np.array_equal(np.signbit(got), np.signbit(exp))
    i = np.array([0, 1, -1, 2 ** 31 - 1, -2 ** 31, 12345, -7], np.int32)
    fn = lambda v: v * 0.5 + 0.1  # noqa: E731
    exp, _ = O.eval_expr(trace(fn, 1), [i])
    assert np.array_equal(ph.NArray.from_numpy(i).process(fn).to_narr().to_numpy(), exp)


def test_int32_division_by_constants(ph):
    """Int32 // and % by a constant: 0 raises DivisionByZeroError, -1 raises
    OverflowError on Int32::MIN // -1, any other constant cannot raise (the
    device then skips the flag read-back); values against the oracle."""
    from crystal import trace
    h = np.array([5, -7, -2 ** 31, 2 ** 31 - 1, 0, 13, -13, 1], np.int32)
    d = ph.NArray.from_numpy(h)
    with pytest.raises(ph.DivisionByZeroError):
        d.map(lambda v: v // 0)
    with pytest.raises(ph.DivisionByZeroError):
        d.map(lambda v: v % 0)
    with pytest.raises(ph.OverflowError):
        d.map(lambda v: v // -1)
    for fn in [lambda v: v % -1, lambda v: v // 3, lambda v: v % 3, lambda v: v // -5, lambda v: v % -5,
               lambda v: (v // 7) % 4, lambda v: v // (2 ** 31 - 1), lambda v: v % -2 ** 31]:
        exp, fl = O.eval_expr(trace(fn), [h])
        assert fl == 0
        assert np.array_equal(d.map(fn).to_numpy(), exp)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_heat_fused_first_operation(ph, dt):
    """Runs of 32 steps or more compute uN - 4u as one fused multiply-add
    once max|u| shows that 4u cannot overflow; the result must stay
    bit-identical to the oracle's separate multiply and subtract, on grids
    that take the clean, edge and masked sweeps."""
    rng = np.random.default_rng(71)
    for ny, nx, steps in [(200, 249, 40), (300, 1000, 33), (65, 257, 35), (1000, 2048, 40), (7, 1, 32)]:
        u0 = (rng.standard_normal((ny, nx)) * 1e3).astype(dt)
        u = ph.NArray.from_numpy(u0)
        e = ph.heat2d(u, 0.2, steps, every=8)
        ref, eref = O.heat2d(u0, 0.2, steps, every=8)
        assert np.array_equal(u.to_numpy(), ref), (ny, nx, steps)
        assert np.allclose(e, eref, rtol=1e-13, atol=1e-300), (ny, nx, steps)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_heat_unfused_when_4u_may_overflow(ph, dt):
    """Grids whose 4u can overflow (values near the largest finite value),
    that hold an infinity or a NaN, or that step with r > 1/4 (values may
    grow) keep the separate multiply and subtract: bit-identical to the
    oracle, infinities and NaN included."""
    big = np.finfo(dt).max
    rng = np.random.default_rng(73)
    cases = []
    u0 = rng.random((64, 300)).astype(dt)
    u0[10, 20] = big * 0.9
    u0[40, 100:104] = -big * 0.6
    cases.append((u0, 0.2))
    u1 = rng.random((64, 300)).astype(dt)
    u1[5, 5] = np.inf
    u1[50, 250] = np.nan
    cases.append((u1, 0.2))
    cases.append((rng.random((64, 300)).astype(dt), 0.3))
    for u0, r in cases:
        u = ph.NArray.from_numpy(u0)
        ph.heat2d(u, r, 34, every=0)
        ref, _ = O.heat2d(u0, r, 34, every=0)
        got = u.to_numpy()
        assert np.array_equal(got, ref, equal_nan=True), r
        # every number bit for bit, zeros' signs included; a NaN's sign is
        # not fixed by IEEE 754 arithmetic (and the Float32 form's fused
        # multiply-add propagates NaN operands differently from the host's)
        num = ~np.isnan(ref)
        assert np.array_equal(got[num].view(np.uint64 if dt == np.float64 else np.uint32),
                              ref[num].view(np.uint64 if dt == np.float64 else np.uint32)), r


def test_heat_fp32_twelve_step_passes(ph):
    """Float32 runs of passes of twelve steps (three halo lanes): the fused
    first operation (runs of 32 steps or more), the flush sweeps (rows in
    whole vectors, a wavefront wide or more) and the masked sweep (ragged or
    narrow grids), bit-exact against the oracle's Float32 formula (u + d*r
    rounded once)."""
    rng = np.random.default_rng(79)
    for ny, nx in [(65, 257), (200, 249), (130, 5), (300, 1000), (1000, 2048), (97, 260), (40, 256)]:
        for steps, every in [(36, 0), (48, 24), (37, 0)]:
            u0 = (rng.standard_normal((ny, nx)) * 10).astype(np.float32)
            u = ph.NArray.from_numpy(u0)
            e = ph.heat2d(u, 0.2, steps, every=every)
            ref, eref = O.heat2d(u0, 0.2, steps, every=every)
            assert np.array_equal(u.to_numpy(), ref), (ny, nx, steps, every)
            if every:
                assert np.allclose(e, eref, rtol=1e-13, atol=1e-300), (ny, nx, steps)
