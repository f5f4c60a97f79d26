"""Enumerable#max (`include Enumerable`, src/multi_indexable.cr:18) keeps
the first element and replaces it only by a greater one: a NaN in first
place is the result, a NaN later is skipped, and of equal zeros the first
one's sign survives.  The device's maxima (full, rows, columns, through
reduce2d_all, slices_max and max, on direct and transposed views) are
compared against the oracle's sequential `>` loop: NaN where it is NaN,
otherwise bit for bit (so the sign of a zero counts)."""
import numpy as np
import pytest

import oracle_api as O
from crystal import trace

pytestmark = pytest.mark.gpu


def _same(got, exp):
    got = np.asarray(got, dtype=np.float64)
    exp = np.asarray(exp, dtype=np.float64)
    nan = np.isnan(exp)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint64), exp[~nan].view(np.uint64))


def _planted(rng, rows, cols):
    """Normal data with NaNs as first elements of the full view, of some
    rows and columns, and mid-slice; rows and columns whose maximum is a
    zero, reached first as -0.0 or as +0.0."""
    x = rng.standard_normal((rows, cols))
    x[0, 0] = np.nan
    x[5, 0] = np.nan
    x[0, 7] = np.nan
    x[9, 100 % cols] = np.nan
    x[30:40] = -np.abs(x[30:40]) - 0.5
    x[31, 3], x[31, 9] = -0.0, 0.0
    x[32, 4], x[32, 6] = 0.0, -0.0
    x[33, cols - 1] = -0.0
    x[:, 50:60] = -np.abs(x[:, 50:60]) - 0.5
    x[11, 51], x[12, 51] = -0.0, 0.0
    x[13, 52], x[200 % rows, 52] = 0.0, -0.0
    x[rows - 1, 53] = -0.0
    return x


def _check_all(pv, f):
    res = pv.reduce2d_all()
    _same([res["full_max"]], [O.max_seq(f)])
    _same(res["row_max"], O.slices_reduce(f, 0, 1))
    _same(res["col_max"], O.slices_reduce(f, 1, 1))
    _same(pv.slices_max(0), O.slices_reduce(f, 0, 1))
    _same(pv.slices_max(1), O.slices_reduce(f, 1, 1))
    _same([pv.max()], [O.max_seq(f)])


@pytest.mark.parametrize("dt", ["F64", "F32"])
def test_max_nan_and_signed_zeros(ph, dt):
    rng = np.random.default_rng(91)
    for rows, cols in [(300, 500), (64, 1027)]:
        h = _planted(rng, rows, cols)
        if dt == "F32":
            h = h.astype(np.float32)
        x = ph.NArray.from_numpy(h)
        f = h.astype(np.float64) * 2.0
        _check_all(x.process(lambda v: v * 2.0), f)
        # the transposed view runs in the storage's orientation
        t = x.view().permute(1, 0).process(lambda v: v * 2.0)
        _check_all(t, np.ascontiguousarray(f.T))


def test_full_max_first_zero(ph):
    """All values <= 0: the full maximum is the first zero in lexicographic
    order, found from each row's first zero (and, for a transposed view,
    from the storage's column order)."""
    rng = np.random.default_rng(93)
    for first_sign, shape in [(-0.0, (200, 300)), (0.0, (200, 300)), (-0.0, (17, 2000)), (0.0, (1000, 33))]:
        h = -np.abs(rng.standard_normal(shape)) - 0.25
        r, c = shape[0] // 3, shape[1] // 2
        h[r, c] = first_sign
        h[r, c + 1 :: 7] = -first_sign  # later zeros of the other sign
        h[r + 1 :: 5, 0] = -first_sign
        x = ph.NArray.from_numpy(h)
        got = x.process(lambda v: v * 1.0).max()
        assert got == 0.0 and np.signbit(got) == np.signbit(first_sign)
        _check_all(x.process(lambda v: v * 1.0), h)
        t = x.view().permute(1, 0).process(lambda v: v * 1.0)
        _check_all(t, np.ascontiguousarray(h.T))


def test_max_of_rank3_slices(ph):
    """slices(k).map(&.max) over a rank-3 array for each k (the middle axis
    goes through a gathered copy)."""
    rng = np.random.default_rng(95)
    h = rng.standard_normal((6, 40, 50))
    h[:, 0, 0] = np.nan
    h[2] = -np.abs(h[2]) - 1
    h[2, 3, 4], h[2, 5, 1] = -0.0, 0.0
    x = ph.NArray.from_numpy(h)
    pv = x.process(lambda v: v * 1.0)
    for k in range(3):
        _same(pv.slices_max(k), O.slices_reduce(h, k, 1))
    _same([pv.max()], [O.max_seq(h)])


@pytest.mark.parametrize("dt", ["F64", "F32"])
def test_signalling_nan_inputs(ph, dt):
    """A signalling NaN in the data (arithmetic never makes one) is skipped
    like any NaN that is not first.  The blocks cover plain loads,
    identities the compiler folds into a pass-through (x * 1.0, x - 0.0),
    whose maxima must re-quiet their operands, and arithmetic (x * 2.0,
    (x + 4) ** 2 * 0.5 - x), whose maxima take the instruction as is.  Each
    signalling NaN sits after its column's and its row's largest element,
    so a maximum that let it through would lose those."""
    rng = np.random.default_rng(97)
    npdt, ubits, snan = ((np.float64, np.uint64, 0x7FF0000000000001) if dt == "F64"
                         else (np.float32, np.uint32, 0x7F800001))
    h = rng.standard_normal((300, 500)).astype(npdt)
    for r, c in [(5, 7), (120, 33), (299, 499), (0, 3)]:
        h[r - 3 if r >= 3 else r + 1, c] = 50 + r
        h[r, c - 2] = 60 + r
        h.view(ubits)[r, c] = ubits(snan)
    h.view(ubits)[100, 0] = ubits(snan)  # first element of row 100
    x = ph.NArray.from_numpy(h)
    with np.errstate(invalid="ignore"):  # widening quiets the signalling NaNs
        f64 = h.astype(np.float64)
    blocks = [("load", None), ("x * 1.0", lambda v: v * 1.0), ("x - 0.0", lambda v: v - 0.0),
              ("x * 2.0", lambda v: v * 2.0), ("C3", lambda v: (v + 4) ** 2 * 0.5 - v)]
    for name, fn in blocks:
        print(f"signalling NaN, {dt}, block {name}", flush=True)
        pv = x.view() if fn is None else x.process(fn)
        f = f64 if fn is None else O.eval_expr(trace(fn), [h])[0].astype(np.float64)
        _check_all(pv, f)
