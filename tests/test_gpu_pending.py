"""Operator results left pending (Storage::pending_map): the reference's
operators are eager (src/multi_indexable.cr:931-985 build each result with
map_with), so `(a + b.tile(n, 1)) * c` is two maps there.  Here a result
whose program cannot raise is left pending and fused into the next
operator's kernel: one launch, the same roundings (bit-exact against the
oracle's two eager steps).  A pending result is what the operator would have
computed at once: writes to its inputs run it first, a map never fuses a
pending operand that reads its own destination, programs that can raise
(checked Int32) run at once, and chains too long for one kernel split."""
import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu


def test_broadcast_operator_form_is_one_kernel(ph):
    rng = np.random.default_rng(101)
    for n, m in [(1024, 1024), (257, 1030)]:
        an, bn, cn = rng.uniform(-1, 1, (n, m)), rng.uniform(-1, 1, (1, m)), rng.uniform(-1, 1, (n, m))
        a, b, c = (ph.NArray.from_numpy(v) for v in (an, bn, cn))
        ph.sync()
        k0 = ph.launch_count()
        r = (a + b.tile(n, 1)) * c
        got = r.to_numpy()
        assert ph.launch_count() - k0 == 1, ph.launch_count() - k0
        t, _ = O.eval_expr("(add x0 x1)", [an, np.tile(bn, (n, 1))])
        exp, _ = O.eval_expr("(mul x0 x1)", [t, cn])
        assert np.array_equal(got, exp)


def test_pending_result_survives_writes_to_its_inputs(ph):
    """c = a + 1 stays what it was when the operator ran, whatever is written
    to `a` afterwards: element writes, chunk writes, in-place maps."""
    h = np.arange(600, dtype=np.float64).reshape(20, 30)
    a = ph.NArray.from_numpy(h)
    c = a + 1.0
    a.set_element([0, 0], -5.0)
    assert np.array_equal(c.to_numpy(), h + 1.0)
    c2 = a * 2.0
    a[["1..3", ".."]] = 0.5
    h2 = h.copy()
    h2[0, 0] = -5.0
    assert np.array_equal(c2.to_numpy(), h2 * 2.0)
    c3 = a - 3.0
    a.map_(lambda v: v * 0.0)
    h3 = h2.copy()
    h3[1:4] = 0.5
    assert np.array_equal(c3.to_numpy(), h3 - 3.0)
    assert np.array_equal(a.to_numpy(), h3 * 0.0)


def test_pending_operand_reading_the_destination_stays_a_snapshot(ph):
    """t = z.permute(1, 0) + 0 is computed before z[mask] = t writes z (the
    reference evaluates the operator first), so the mask assignment sees t's
    values, not z's as it is being written."""
    rng = np.random.default_rng(103)
    h = rng.integers(-100, 100, (96, 96)).astype(np.float64)
    mh = rng.random((96, 96)) < 0.5
    z = ph.NArray.from_numpy(h)
    t = z.view().permute(1, 0) + 0.0
    z[ph.NArray.from_numpy(mh)] = t
    assert np.array_equal(z.to_numpy(), np.where(mh, h.T, h))


def test_checked_int32_operators_raise_at_once(ph):
    a = ph.NArray.from_numpy(np.full((10,), 2 ** 30, np.int32))
    with pytest.raises(ph.OverflowError):
        a + a  # the reference raises in the operator itself
    w = a.wadd(a)  # wrapping: cannot raise, may be left pending
    assert np.array_equal(w.to_numpy(), np.full(10, -2 ** 31, np.int32))


def test_long_chains_split_into_kernels(ph):
    """A chain of 80 operators is more than one fused program holds: it runs
    in pieces, bit-exact."""
    h = np.linspace(-1, 1, 5000)
    x = ph.NArray.from_numpy(h)
    ref = h.copy()
    for k in range(80):
        x = x * 1.0001 + 0.25 if k % 2 else x - 0.125
        ref = ref * 1.0001 + 0.25 if k % 2 else ref - 0.125
    assert np.array_equal(x.to_numpy(), ref)


def test_unused_pending_result_costs_nothing(ph):
    a = ph.NArray.from_numpy(np.ones((1000,)))
    ph.sync()
    k0 = ph.launch_count()
    for _ in range(10):
        _ = a * 3.0 + 1.0
    assert ph.launch_count() == k0
    s = (a * 3.0 + 1.0).sum()  # a reduction runs the pending map first
    assert s == 4000.0


def test_mask_used_many_times_is_computed_once_more(ph):
    """A pending Bool mask computed from Int32 values reads 4 bytes per
    element to give 1: it is spliced into its first consumer, and the second
    consumer runs it into its own buffer (two launches), so the later ones
    read the stored mask (one launch each).  A pending result that reads no
    more than it holds (a broadcast sum) is spliced every time.  Every
    result bit-exact against the oracle."""
    rng = np.random.default_rng(107)
    xh = rng.integers(-2 ** 31, 2 ** 31, (300, 200), dtype=np.int64).astype(np.int32)
    zh = rng.integers(-2 ** 31, 2 ** 31, (300, 200), dtype=np.int64).astype(np.int32)
    x = ph.NArray.from_numpy(xh)
    mask = x.process(lambda v: (v & 1) == 0).to_narr()
    mh = (xh & 1) == 0
    launches = []
    for k in range(4):
        z = ph.NArray.from_numpy(zh)
        ph.sync()
        k0 = ph.launch_count()
        z[mask] = 7 + k
        z.materialize()
        launches.append(ph.launch_count() - k0)
        assert np.array_equal(z.to_numpy(), O.set_mask(zh.copy(), mh, 7 + k)), k
    assert launches == [1, 2, 1, 1], launches
    an, bn = rng.uniform(-1, 1, (300, 200)), rng.uniform(-1, 1, (1, 200))
    t = ph.NArray.from_numpy(an) + ph.NArray.from_numpy(bn).tile(300, 1)
    exp, _ = O.eval_expr("(add x0 x1)", [an, np.tile(bn, (300, 1))])
    for k in range(3):
        ph.sync()
        k0 = ph.launch_count()
        got = (t * float(k + 2)).to_numpy()
        assert ph.launch_count() - k0 == 1
        assert np.array_equal(got, exp * float(k + 2)), k
