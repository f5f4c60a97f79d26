"""NArray#map_with_index / map_with_index! / map_with_coord / map_with_coord!
(src/n_array.cr:605-650) and MultiIndexable#map_with_coord over views
(src/multi_indexable.cr:708-712) on the device, against the reference specs
(spec/n_array_spec.cr:451-518) and the oracle.

The oracle side is map_with over the element and extra Int32 operands:
np.arange for the lexicographic index and np.indices for the coordinate.
The device derives both from the element's position inside the map kernel,
so the two sides share no code."""
import json
import os

import numpy as np
import pytest

import oracle_api as O
from crystal import trace

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_spec_vectors.json")))


def _flat(a):
    return a.to_numpy().reshape(-1).tolist()


def test_spec_known_answers(ph):
    g = GOLD["narray_maps"]
    stock = ph.NArray.from_numpy(np.arange(6, dtype=np.int32).reshape(2, 3))
    # map_with_coord and map_with_coord! (n_array_spec.cr:451-458, :484-500)
    assert _flat(stock.map_with_coord(lambda el, c: 10 * c[0] + c[1] + el)) == g["with_coord"]
    narr = stock.clone()
    r = narr.map_with_coord_(lambda el, c: 10 * c[0] + c[1] + el)
    assert _flat(narr) == g["with_coord"] and _flat(r) == g["with_coord"]
    # map_with_index and map_with_index! (:460-467, :503-518)
    narr = stock * 2
    assert _flat(narr.map_with_index(lambda el, idx: idx + el)) == g["times3"]
    assert _flat(narr) == [2 * v for v in range(6)]  # the source is not changed
    r = narr.map_with_index_(lambda el, idx: idx + el)
    assert _flat(narr) == g["times3"] and _flat(r) == g["times3"]
    # the map_with_coord doc example (multi_indexable.cr:700-707), coord.sum
    doc = ph.NArray.from_numpy(np.array(g["doc_coord_values"], np.int32).reshape(g["doc_coord_shape"]))
    assert _flat(doc.map_with_coord(lambda el, c: el + sum(c))) == g["doc_coord_sum"]
    # the index is yielded third to a three-parameter block (n_array.cr:614-618)
    assert _flat(stock.map_with_coord(lambda el, c, idx: idx * 100 + c[1])) == [0, 101, 202, 300, 401, 502]


def test_coordinate_order_known_answers(ph):
    """Element and coordinate pairs in lexicographic order, and in
    colexicographic order as the lexicographic order of the transposed view
    with each coordinate read back to front (elem_coord_iterator_spec.cr:
    10-22)."""
    g = GOLD["elem_coord_order"]
    narr = ph.NArray.from_numpy(np.array(g["values"], np.int32))
    enc = lambda el, c: el * 100 + c[0] * 10 + c[1]  # noqa: E731
    assert _flat(narr.map_with_coord(enc)) == [e * 100 + c[0] * 10 + c[1] for e, c in g["lex"]]
    got = _flat(narr.permute(1, 0).map_with_coord(lambda el, c: enc(el, c[::-1])))
    assert got == [e * 100 + c[0] * 10 + c[1] for e, c in g["colex"]]


def _oracle(fn, vals, bf16=False):
    """fn(el, coord, idx) over the host values, by the oracle's zip evaluator."""
    shape = vals.shape
    idx = np.arange(vals.size, dtype=np.int32).reshape(shape)
    coords = list(np.indices(shape, dtype=np.int32))
    return O.eval_expr(trace(lambda el, k, *c: fn(el, list(c), k), 2 + len(shape)), [vals, idx] + coords,
                       bf16=(0,) if bf16 else ())


BLOCKS = {
    "F64": [lambda el, c, k: el * 0.5 + k, lambda el, c, k: el - c[0] * 3 + c[-1]],
    "F32": [lambda el, c, k: el * k, lambda el, c, k: (el + c[-1]) / (c[0] + 1)],
    "BF16": [lambda el, c, k: el + k, lambda el, c, k: el * c[-1] > c[0]],
    "I32": [lambda el, c, k: k + el, lambda el, c, k: (el ^ k) & 1023 + c[0] * c[-1],
            lambda el, c, k: k.cmp(el) * c[-1]],
}


@pytest.mark.parametrize("storage", ["F64", "F32", "BF16", "I32"])
def test_random_shapes_against_oracle(ph, storage):
    """Row-tiled and flat kernel forms, extent-1 axes that the layout folds
    away, ranks 1 to 4, each element type; bit-exact."""
    rng = np.random.default_rng(41)
    for shape in [(1,), (7,), (3000,), (3, 5, 7), (2, 1, 1030), (300, 1000), (4, 3, 2, 260), (100, 100),
                  (1, 1), (517, 3)]:
        if storage == "I32":
            a = rng.integers(-10000, 10000, shape).astype(np.int32)
        else:
            a = rng.standard_normal(shape).astype(np.float64 if storage == "F64" else np.float32)
        if storage == "BF16":
            bits = (a.view(np.uint32) >> 16).astype(np.uint16)
            d, host = ph.NArray.from_numpy(bits, ph.DType.BF16), bits
        else:
            d, host = ph.NArray.from_numpy(a), a
        for fn in BLOCKS[storage]:
            exp, fl = _oracle(fn, host, bf16=storage == "BF16")
            assert fl == 0
            got = d.map_with_coord(fn).to_numpy()
            assert got.dtype == exp.dtype and np.array_equal(got, exp), (shape, storage)
            # the index in every coordinate's place: products of large indices
            # overflow Int32 and must raise as the oracle flags them
            blk = lambda el, k: fn(el, [k] * len(shape), k)  # noqa: E731
            exp_i, fl_i = O.eval_expr(trace(blk, 2), [host, np.arange(host.size, dtype=np.int32).reshape(shape)],
                                      bf16=(0,) if storage == "BF16" else ())
            if fl_i:
                with pytest.raises(ph.OverflowError):
                    d.map_with_index(blk)
            else:
                assert np.array_equal(d.map_with_index(blk).to_numpy(), exp_i), (shape, storage)
            if storage != "BF16" and exp.dtype == a.dtype:
                e = ph.NArray.from_numpy(a)
                e.map_with_coord_(fn)
                assert np.array_equal(e.to_numpy(), exp), (shape, storage)


def test_views_and_procviews(ph):
    """map_with_coord over a strided, reversed, permuted, reshaped or lazily
    mapped view sees the coordinate in the view's own shape."""
    rng = np.random.default_rng(43)
    a = rng.standard_normal((60, 700))
    d = ph.NArray.from_numpy(a)
    fn = lambda el, c, k: el * 2.0 + c[0] * 1000 - c[1] + k * 0.25  # noqa: E731
    cases = [
        (d.view("3..2..", "-1..0"), a[3::2, ::-1]),
        (d.permute(1, 0), a.T),
        (d.reshape(700, 60), a.reshape(700, 60)),
        (d.process(lambda v: v - 1.5), a - 1.5),
    ]
    for v, host in cases:
        exp, _ = _oracle(fn, np.ascontiguousarray(host))
        assert np.array_equal(v.map_with_coord(fn).to_numpy(), exp)


def test_index_overflow_empty_and_size_limit(ph):
    big = 2 ** 31 - 20
    a = ph.NArray.fill([3, 10], big)
    with pytest.raises(ph.OverflowError):  # checked Int32 `+` (element 20 reaches 2^31)
        a.map_with_index(lambda el, k: el + k)
    w = a.map_with_index(lambda el, k: el.wadd(k)).to_numpy().reshape(-1)
    exp, _ = O.eval_expr(trace(lambda el, k: el.wadd(k), 2), [np.full(30, big, np.int32), np.arange(30, dtype=np.int32)])
    assert np.array_equal(w, exp)
    e = ph.NArray.fill([0, 5], 1.5)
    assert e.map_with_index(lambda el, k: el + k).to_numpy().shape == (0, 5)
    assert e.map_with_coord(lambda el, c: el + c[1]).to_numpy().shape == (0, 5)
    # Crystal's Int32 index cannot number more than 2^31 - 1 elements
    huge = ph.NArray.fill([2 ** 31], False)
    with pytest.raises(ph.OverflowError):
        huge.map_with_index(lambda el, k: el | (k > 5))
    del huge
