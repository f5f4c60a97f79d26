"""NArray#[]=(mask, value) and (mask, source) (src/n_array.cr:510-536) on the
device at sizes beyond the spec's 2 x 3, against the oracle's mask loop:
every element type, masks that are NArrays, comparisons and lazy views, and
the shape errors the reference raises."""
import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu


def _pair(ph, rng, shape, kind):
    if kind == "i32":
        h = rng.integers(-2 ** 31, 2 ** 31, shape, dtype=np.int64).astype(np.int32)
        return ph.NArray.from_numpy(h), h
    if kind == "f64":
        h = rng.standard_normal(shape)
        return ph.NArray.from_numpy(h), h
    if kind == "bf16":
        h = rng.integers(0, 2 ** 16, shape, dtype=np.int64).astype(np.uint16)
        return ph.NArray.from_numpy(h, ph.DType.BF16), h
    h = rng.random(shape) < 0.5
    return ph.NArray.from_numpy(h), h


@pytest.mark.parametrize("kind", ["i32", "f64", "bool"])
def test_masks_against_oracle(ph, kind):
    rng = np.random.default_rng(61)
    scalar = {"i32": -7, "f64": 2.5, "bool": True}[kind]
    for shape in [(1,), (1000,), (37, 41), (300, 1027), (3, 5, 700)]:
        d, h = _pair(ph, rng, shape, kind)
        mh = rng.random(shape) < 0.3
        mask = ph.NArray.from_numpy(mh)
        e = d.clone()
        e[mask] = scalar
        assert np.array_equal(e.to_numpy(), O.set_mask(h.copy(), mh, scalar)), (shape, kind)
        s, sh = _pair(ph, rng, shape, kind)
        e = d.clone()
        e[mask] = s
        assert np.array_equal(e.to_numpy(), O.set_mask(h.copy(), mh, sh)), (shape, kind)


def test_mask_from_comparison_and_view(ph):
    """The mask is a comparison result or a lazy view; the source a view."""
    rng = np.random.default_rng(67)
    h = rng.integers(-1000, 1000, (400, 300)).astype(np.int32)
    d = ph.NArray.from_numpy(h)
    e = d.clone()
    e[d.process(lambda v: v > 0)] = 0
    assert np.array_equal(e.to_numpy(), O.set_mask(h.copy(), h > 0, 0))
    src = d.permute(1, 0).reshape(400, 300)
    sh = np.ascontiguousarray(h.T).reshape(400, 300)
    e = d.clone()
    e[d.process(lambda v: (v & 3) == 1)] = src
    assert np.array_equal(e.to_numpy(), O.set_mask(h.copy(), (h & 3) == 1, sh))


def test_mask_errors(ph):
    d = ph.NArray.fill([4, 5], 1)
    with pytest.raises(ph.DimensionError, match="mask shape does not match"):
        d[ph.NArray.fill([5, 4], True)] = 3
    with pytest.raises(ph.ShapeError):
        d[ph.NArray.fill([4, 5], True)] = ph.NArray.fill([5, 4], 2)


def test_mask_source_aliasing_the_destination(ph):
    """z[mask] = (a view of z): the reference's loop (n_array.cr:523-536)
    reads each source element after the writes that precede it in
    lexicographic order, so the result differs from reading a snapshot.
    The expected values replay that loop element by element on small
    arrays."""
    rng = np.random.default_rng(71)
    differs = 0
    for shape, kind in [((6, 6), "T"), ((5, 7), "rev"), ((40, 40), "T"), ((3, 4, 5), "rev"), ((9, 9), "T")]:
        h = rng.integers(-1000, 1000, shape).astype(np.int32)
        mh = rng.random(shape) < 0.5
        z = ph.NArray.from_numpy(h)
        if kind == "T":
            src = z.view().permute(1, 0)
            idx = lambda c: (c[1], c[0])  # noqa: E731
        else:
            src = z.view().reverse()
            idx = lambda c, s=shape: tuple(n - 1 - ci for ci, n in zip(c, s))  # noqa: E731
        z[ph.NArray.from_numpy(mh)] = src
        exp = h.copy()
        for c in np.ndindex(*shape):
            if mh[c]:
                exp[c] = exp[idx(c)]
        assert np.array_equal(z.to_numpy(), exp), (shape, kind)
        snap = h.copy()
        for c in np.ndindex(*shape):
            if mh[c]:
                snap[c] = h[idx(c)]
        differs += int(not np.array_equal(snap, exp))
    assert differs >= 3  # the cases do tell the two readings apart
