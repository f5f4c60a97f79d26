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
