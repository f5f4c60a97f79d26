"""NArray.concatenate (device strided copies, checked bit for bit against the
oracle's restatement of concatenate_to_slice) and JSON / YAML serialisation
of device arrays (the reference's spec strings; round trips)."""
import json
import os

import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_spec_vectors.json")))


def _rand(rng, shape, dt):
    if dt == np.float64 or dt == np.float32:
        return rng.standard_normal(shape).astype(dt)
    if dt == np.bool_:
        return rng.random(shape) < 0.5
    return rng.integers(-2 ** 31, 2 ** 31, shape, dtype=np.int64).astype(np.int32)


@pytest.mark.parametrize("dt", [np.int32, np.float32, np.float64, np.bool_])
def test_concatenate_random_vs_oracle(ph, dt):
    rng = np.random.default_rng(21)
    for _ in range(40):
        rank = int(rng.integers(1, 5))
        base = [int(v) for v in rng.integers(1, 6, rank)]
        axis = int(rng.integers(0, rank))
        arrs = []
        for _ in range(int(rng.integers(1, 5))):
            s = list(base)
            s[axis] = int(rng.integers(0, 6))
            arrs.append(_rand(rng, s, dt))
        dev = [ph.NArray.from_numpy(a) for a in arrs]
        got = ph.NArray.concatenate_all(*dev, axis=axis).to_numpy()
        want = O.concatenate([a.astype(np.uint8) if dt == np.bool_ else a for a in arrs], axis)
        if dt == np.bool_:
            want = want.astype(np.bool_)
        assert got.shape == want.shape and np.array_equal(got, want), (base, axis)


def test_concatenate_views_procviews_and_in_place(ph):
    rng = np.random.default_rng(5)
    a = rng.integers(-100, 100, (6, 8)).astype(np.int32)
    b = rng.integers(-100, 100, (6, 5)).astype(np.int32)
    da, db = ph.NArray.from_numpy(a), ph.NArray.from_numpy(b)
    v = da.view("..", "..2..")  # [6, 4]
    pv = db.process(lambda x: x * 3 + 1)
    got = ph.NArray.concatenate_all(v, pv, da, axis=1).to_numpy()
    want = O.concatenate([a[:, ::2], b * 3 + 1, a], 1)
    assert np.array_equal(got, want)
    # instance form, negative axis (shapes equal on every axis), in place, push
    c = ph.NArray.from_numpy(a)
    assert np.array_equal(c.concatenate(c, axis=-1).to_numpy(), O.concatenate([a, a], 1))
    c.concatenate_(ph.NArray.from_numpy(a), axis=0)
    assert np.array_equal(c.to_numpy(), O.concatenate([a, a], 0))
    c.push(ph.NArray.from_numpy(a[:2]))
    assert np.array_equal(c.to_numpy(), O.concatenate([a, a, a[:2]], 0))


def test_concatenate_errors(ph):
    a = ph.NArray.from_numpy(np.zeros((2, 3), np.int32))
    b = ph.NArray.from_numpy(np.zeros((3, 3), np.int32))
    with pytest.raises(ph.DimensionError, match="Cannot concatenate these arrays along axis 1: shapes do not match"):
        ph.NArray.concatenate_all(a, b, axis=1)
    with pytest.raises(ph.DimensionError):
        ph.NArray.concatenate_all(a, b, axis=-1)  # a negative axis exempts no dimension
    with pytest.raises(ph.ArgumentError):
        ph.NArray.concatenate_all(a, ph.NArray.from_numpy(np.zeros((2, 3))), axis=0)


def test_concatenate_large(ph):
    n = 4096
    parts = [ph.NArray.random([n, n], ph.DType.F32, seed=30 + k) for k in range(4)]
    got = ph.NArray.concatenate_all(*parts, axis=1)
    hs = [p.to_numpy() for p in parts]
    assert np.array_equal(got.to_numpy(), O.concatenate(hs, 1))


# ------------------------------------------------------------ serialisation
def test_json_yaml_spec_strings(ph):
    g = GOLD["narray_serialisation"]
    stock = ph.NArray.from_numpy(np.array(g["stock"], np.int32).reshape(g["stock_shape"]))
    empty = ph.NArray.fill([0], 0)
    assert stock.to_json() == g["to_json"]
    assert empty.to_json() == g["empty_to_json"]
    assert stock.to_yaml() == g["to_yaml"]
    assert empty.to_yaml() == g["empty_to_yaml"]
    assert ph.NArray.from_json(g["to_json"], ph.DType.I32).equals(stock)
    e = ph.NArray.from_json(g["empty_to_json"], ph.DType.I32)
    assert list(e.shape) == [0]
    assert ph.NArray.from_yaml(g["to_yaml"], ph.DType.I32).equals(stock)
    assert list(ph.NArray.from_yaml(g["empty_to_yaml"], ph.DType.I32).shape) == [0]


def test_json_yaml_errors(ph):
    with pytest.raises(ph.ParseError, match="'shape' and/or 'elements' were missing"):
        ph.NArray.from_json('{"shape":[2]}', ph.DType.I32)
    with pytest.raises(ph.ParseError, match=r"wrong number of elements for shape \[2, 2\]"):
        ph.NArray.from_json('{"shape":[2,2],"elements":[1,2,3]}', ph.DType.I32)
    with pytest.raises(ph.ParseError):
        ph.NArray.from_json('{"shape":[2],"elements":[1.5,2]}', ph.DType.I32)
    with pytest.raises(ph.ParseError, match="'shape' and/or 'elements' were missing"):
        ph.NArray.from_yaml("---\nshape: [2]\n", ph.DType.I32)
    with pytest.raises(ph.ParseError, match=r"wrong number of elements for shape \[3\]"):
        ph.NArray.from_yaml("---\nshape: [3]\nelements: [1, 2]\n", ph.DType.I32)
    with pytest.raises(ph.ParseError):
        ph.NArray.from_json(ph.NArray.from_numpy(np.array([np.nan])).to_json(), ph.DType.F64)


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.int32, np.bool_])
def test_json_yaml_round_trip(ph, dt):
    rng = np.random.default_rng(9)
    a = _rand(rng, (7, 33), dt)
    if dt in (np.float64, np.float32):
        a = (a * 10.0 ** rng.integers(-20, 20, a.shape)).astype(dt)
    d = ph.NArray.from_numpy(a)
    D = {np.float64: ph.DType.F64, np.float32: ph.DType.F32, np.int32: ph.DType.I32, np.bool_: ph.DType.Bool}[dt]
    j = d.to_json()
    assert json.loads(j)["shape"] == [7, 33]
    for text, back in [(j, ph.NArray.from_json(j, D)), (d.to_yaml(), ph.NArray.from_yaml(d.to_yaml(), D))]:
        b = back.to_numpy()
        assert b.dtype == a.dtype and np.array_equal(b, a), text[:200]
    y = d.to_yaml()
    assert all(len(line) <= 80 + 40 for line in y.splitlines())  # flow sequences wrap near column 80


def test_json_of_many_elements_in_chunks(ph):
    """Documents long enough to be written and read in chunks on several
    threads: the text is Python's own JSON of the same values, in order, and
    reads back bit for bit; the first non-finite or invalid element in
    order is the one reported, though later chunks hold others."""
    rng = np.random.default_rng(10)
    n = 300007
    a = rng.standard_normal(n) * 10.0 ** rng.integers(-30, 30, n)
    d = ph.NArray.from_numpy(a)
    j = d.to_json()
    doc = json.loads(j)
    assert doc["shape"] == [n] and np.array_equal(np.array(doc["elements"]).view(np.uint64), a.view(np.uint64))
    assert np.array_equal(ph.NArray.from_json(j, ph.DType.F64).to_numpy().view(np.uint64), a.view(np.uint64))
    ai = rng.integers(-2 ** 31, 2 ** 31, n, dtype=np.int64).astype(np.int32)
    ji = ph.NArray.from_numpy(ai).to_json()
    assert json.loads(ji)["elements"] == ai.tolist()
    b = a.copy()
    b[n - 5], b[200003], b[150001] = np.nan, np.inf, -np.inf
    with pytest.raises(ph.ParseError, match="^-Infinity not allowed in JSON"):
        ph.NArray.from_numpy(b).to_json()
    parts = j.split(",")
    parts[250000] = "1.5x"
    parts[120000] = "2.5y"
    with pytest.raises(ph.ParseError, match="invalid Float64 element '2.5y'"):
        ph.NArray.from_json(",".join(parts), ph.DType.F64)
