"""Reductions over a few long series: rank-1 arrays, keep-axis-0 sums of a
handful of very wide rows, per-column results of tall arrays.
Enumerable#sum and #max over `each` (src/multi_indexable.cr:18, :656-658)
and `slices(k).map(&.sum)` (:742-786) are checked against the oracle's
sequential loops.  Int32 sums are exact, and so is the running total that
raises OverflowError.  Float sums are within 1e-12 of the exact sum, and
maxima are bit for bit.  A whole contiguous view is re-split into rows of
about 16384 elements.  Long series without such a split have their tile
states pre-folded in ordered runs of 64.  The timing case shows a rank-1
sum of 2^27 elements streams: it took about 30 ms while one thread folded
its 131072 tile states."""
import time

import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu


def _prime_at_least(n):
    def is_prime(k):
        if k % 2 == 0:
            return False
        d = 3
        while d * d <= k:
            if k % d == 0:
                return False
            d += 2
        return True
    while not is_prime(n):
        n += 1
    return n


def _same(got, exp):
    got = np.asarray(got, dtype=np.float64)
    exp = np.asarray(exp, dtype=np.float64)
    nan = np.isnan(exp)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint64), exp[~nan].view(np.uint64))


def _close(got, exp, tol=1e-12):
    """Within tol of the exact sums; NaN exactly where they are NaN."""
    got = np.asarray(got, dtype=np.float64)
    exp = np.asarray(exp, dtype=np.float64)
    nan = np.isnan(exp)
    assert np.array_equal(np.isnan(got), nan)
    g, e = got[~nan], exp[~nan]
    assert np.all(np.abs(g - e) <= tol * np.maximum(np.abs(e), 1e-300)), np.max(np.abs(g - e))


def test_rank1_int32_sum_of_2_27_streams(ph):
    n = 1 << 27
    x = ph.NArray.random([n], ph.DType.I32, seed=21, a=-1000, b=1000)
    h = x.to_numpy()
    exp, rc = O.sum_i32(h)
    assert rc == 0
    assert x.sum() == exp
    assert x.max() == int(h.max())
    ph.sync()
    t0 = time.perf_counter()
    for _ in range(5):
        s = x.sum()
    dt = (time.perf_counter() - t0) / 5
    assert s == exp
    print(f"rank-1 Int32 sum of 2^27: {dt * 1e3:.3f} ms per call (with the result's read-back)")
    assert dt < 3e-3


def test_rank1_overflow_of_the_running_total(ph):
    """The running total leaves Int32 at element 1 and comes back: the
    reference's checked `+` raises there, whatever the final total."""
    n = 1 << 26
    h = np.zeros(n, np.int32)
    h[1 << 20], h[(1 << 20) + 1] = 2 ** 30, 2 ** 30
    h[(1 << 25) + 3], h[(1 << 25) + 4] = -(2 ** 30), -(2 ** 30)
    assert O.sum_i32(h)[1] != 0
    x = ph.NArray.from_numpy(h)
    with pytest.raises(OverflowError):
        x.sum()
    h2 = h.copy()
    h2[(1 << 20) + 1] = -(2 ** 30)  # never leaves the range
    assert ph.NArray.from_numpy(h2).sum() == O.sum_i32(h2)[0]


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_prime_length_rank1(ph, dt):
    """A prime element count has no row split: one row of many column tiles,
    whose states are pre-folded."""
    n = _prime_at_least((1 << 24) + 12345)
    rng = np.random.default_rng(22)
    h = rng.standard_normal(n).astype(dt)
    x = ph.NArray.from_numpy(h)
    _close([x.sum(acc="f64")], [O.sum_exact(h)])
    _close([x.sum()], [O.sum_exact(h)])
    _same([x.max()], [O.max_seq(h)])
    f = x.process(lambda v: v * 2.0 - 1.0)
    fh = h.astype(np.float64) * 2.0 - 1.0
    _close([f.sum(acc="f64")], [O.sum_exact(fh)])
    _same([f.max()], [O.max_seq(fh)])


def test_rank1_max_semantics_long(ph):
    """NaN first, NaN later, and the first zero's sign, over 2^24 elements
    (re-split) and a prime count near 2^21 (pre-folded)."""
    rng = np.random.default_rng(23)
    for n in (1 << 24, _prime_at_least((1 << 21) + 7)):
        h = -np.abs(rng.standard_normal(n)) - 0.5
        h[n // 3], h[n // 3 + 5], h[n - 2] = -0.0, 0.0, 0.0
        x = ph.NArray.from_numpy(h)
        got = x.max()
        assert got == 0.0 and np.signbit(got), n
        h[n // 2] = np.nan
        _same([ph.NArray.from_numpy(h).max()], [O.max_seq(h)])
        h[0] = np.nan
        _same([ph.NArray.from_numpy(h).max()], [O.max_seq(h)])


def test_few_wide_rows_per_slice(ph):
    """keep axis 0 of 3 x (2^24 + 3): every row folds about 16000 tile
    states (pre-folded), sums and maxima per row and the whole."""
    rng = np.random.default_rng(24)
    rows, cols = 3, (1 << 24) + 3
    h = rng.standard_normal((rows, cols)).astype(np.float32)
    h[1, 0] = np.nan
    x = ph.NArray.from_numpy(h)
    f = h.astype(np.float64)
    _close(x.slices_sum(0, acc="f64"), [O.sum_exact(f[r]) for r in range(rows)])
    _same(x.slices_max(0), O.slices_reduce(f, 0, 1))
    r = x.reduce2d_all(acc="f64")
    _close(r["row_sum"], [O.sum_exact(f[k]) for k in range(rows)])
    _close([r["full_sum"]], [O.sum_exact(f)])
    _same([r["full_max"]], [O.max_seq(f)])
    _same(r["row_max"], O.slices_reduce(f, 0, 1))
    hi = rng.integers(-1000, 1000, (rows, cols)).astype(np.int32)
    xi = ph.NArray.from_numpy(hi)
    assert list(xi.slices_sum(0)) == [O.sum_i32(hi[k])[0] for k in range(rows)]
    assert xi.sum() == O.sum_i32(hi)[0]


def test_tall_columns_per_slice(ph):
    """slices(1) of a (2^21 + 5) x 3 array: every column folds thousands of
    row-tile states (pre-folded)."""
    rng = np.random.default_rng(25)
    rows, cols = (1 << 21) + 5, 3
    h = rng.standard_normal((rows, cols))
    h[0, 2] = np.nan
    x = ph.NArray.from_numpy(h)
    _close(x.slices_sum(1, acc="f64"), [O.sum_exact(h[:, c]) for c in range(cols)])
    _same(x.slices_max(1), O.slices_reduce(h, 1, 1))
    r = x.reduce2d_all(acc="f64")
    _close(r["col_sum"], [O.sum_exact(h[:, c]) for c in range(cols)])
    _same(r["col_max"], O.slices_reduce(h, 1, 1))
    _close([r["full_sum"]], [O.sum_exact(h)])


def test_zero_maximum_found_far_along_long_series(ph):
    """Series whose maximum is a zero first reached near their end: the
    first zero's sign is the result (Enumerable#max keeps it), found by a
    wavefront scanning 512 elements per round.  A prime-length rank-1
    array (one long row), a column of a tall array and a row of a wide
    one."""
    rng = np.random.default_rng(26)
    n = _prime_at_least((1 << 24) + 99)
    h = -np.abs(rng.standard_normal(n)) - 0.5
    h[n - 3], h[n - 1] = -0.0, 0.0
    x = ph.NArray.from_numpy(h)
    x.max()
    ph.sync()
    t0 = time.perf_counter()
    got = x.max()
    dt = time.perf_counter() - t0
    assert got == 0.0 and np.signbit(got)
    print(f"zero maximum found at {n - 3} of {n}: {dt * 1e3:.1f} ms")
    assert dt < 1.0
    rows, cols = (1 << 21) + 5, 3
    t = -np.abs(rng.standard_normal((rows, cols))) - 0.5
    t[rows - 2, 1], t[rows - 1, 1] = -0.0, 0.0
    t[rows - 1, 2] = 0.0
    xt = ph.NArray.from_numpy(t)
    _same(xt.slices_max(1), O.slices_reduce(t, 1, 1))
    _same(xt.reduce2d_all()["col_max"], O.slices_reduce(t, 1, 1))
    w = -np.abs(rng.standard_normal((3, (1 << 22) + 3))) - 0.5
    w[2, -5], w[2, -2] = 0.0, -0.0
    w[0, -1] = -0.0
    xw = ph.NArray.from_numpy(w)
    _same(xw.slices_max(0), O.slices_reduce(w, 0, 1))
    r = xw.reduce2d_all()
    _same(r["row_max"], O.slices_reduce(w, 0, 1))
    _same([r["full_max"]], [O.max_seq(w)])
