"""General batches (ph_gbatch_*): each record carries its own map program,
reduction program and region, and the whole batch runs as one persistent
launch.  Checked per array against the oracle's restatement of what a host
loop over ph-core's NArrays computes:

    y_k = x_k.map { |v| map(v, s_k) }            bit-exact (O.eval_expr)
    z_k = x_k.view(region).to_narr               bit-exact (O.region, O.gather)
    v.max, v = x_k.map { |v| red(v, s_k) }       bit-exact, sign of zero and
                                                 NaN included (O.max_seq, the
                                                 `>` loop)
    v.sum                                        |gpu - exact| <= n 2^-52 sum|v|
                                                 (FP64 accumulation, O.sum_exact)

Modes 0 (graph replay), 1 (direct launches), 2 (one launch per array) and 3
(each record through ph_map, ph_reduce and ph_strided_copy, which compile a
kernel per distinct shape, so only small batches use it) must agree: y, z
and maxima bit for bit, sums within the same bound."""
import numpy as np
import pytest

import oracle_api as O
from crystal import parse_literal, trace

pytestmark = pytest.mark.gpu


def _lits(rng, shape):
    """A random region literal per axis: a single index (the axis is
    dropped), or first..step..last with either sign of step."""
    out = []
    for d in shape:
        r = rng.random()
        if r < 0.15:
            out.append(str(int(rng.integers(0, d))))
        elif r < 0.3:
            out.append("..")
        else:
            a, b = sorted(int(v) for v in rng.integers(0, d, 2))
            s = int(rng.integers(1, 4))
            out.append(f"{a}..{s}..{b}" if rng.random() < 0.5 else f"{b}..{-s}..{a}")
    return out


def _records(rng, n, nprog_map, nprog_red, max_side=40):
    recs = []
    for _ in range(n):
        rank = int(rng.integers(1, 4))
        shape = [int(rng.integers(1, max_side)) for _ in range(rank)]
        rec = {"shape": shape, "scalar": float(np.float32(rng.uniform(-2, 2)))}
        if nprog_map and rng.random() < 0.7:
            rec["map"] = int(rng.integers(0, nprog_map))
        if nprog_red and rng.random() < 0.7:
            rec["red"] = nprog_map + int(rng.integers(0, nprog_red))
            rec["ops"] = [["sum"], ["max"], ["sum", "max"]][int(rng.integers(0, 3))]
        if rng.random() < 0.7:
            rec["region"] = _lits(rng, shape)
        recs.append(rec)
    return recs


def _check(ph, b, recs, progs, xdt_np, x):
    y = b.y.to_numpy()
    z = b.z.to_numpy()
    sums, maxes = np.array(b.sums()), np.array(b.maxes())
    for k, rec in enumerate(recs):
        shape = rec["shape"]
        n = int(np.prod(shape))
        o = b.offsets_xy[k]
        xk = x[o:o + n].reshape(shape)
        sk = np.full(shape, rec["scalar"], dtype=xdt_np)
        if "map" in rec:
            exp, fl = O.eval_expr(trace(progs[rec["map"]], 2), [xk, sk])
            got = y[o:o + n].reshape(shape)
            assert np.array_equal(got.view(np.uint8), np.ascontiguousarray(exp).view(np.uint8)), ("y", k, rec)
        if "red" in rec:
            v, _ = O.eval_expr(trace(progs[rec["red"]], 2), [xk, sk])
            v = v.astype(np.float64).reshape(-1)
            if "sum" in rec["ops"]:
                ex = O.sum_exact(v)
                tol = n * 2.0 ** -52 * float(np.abs(v).sum())
                assert abs(sums[k] - ex) <= tol, ("sum", k, rec, sums[k], ex)
            if "max" in rec["ops"]:
                em = O.max_seq(v)
                assert np.array_equal(np.float64(maxes[k]).view(np.uint64), np.float64(em).view(np.uint64)) or \
                    (np.isnan(maxes[k]) and np.isnan(em)), ("max", k, rec, maxes[k], em)
        if "region" in rec:
            r = O.region([parse_literal(l) for l in rec["region"]], shape)
            exp = O.gather(np.ascontiguousarray(xk), r).reshape(-1)
            zo = b.offsets_z[k]
            assert np.array_equal(z[zo:zo + exp.size].view(np.uint8), exp.view(np.uint8)), ("z", k, rec)
            assert list(b.z_at(k).shape) == list(r.shape), (k, rec)


def _fill(ph, b, rng, xdt_np):
    n = b.x.size
    if xdt_np == np.int32:
        h = rng.integers(-1000, 1000, n).astype(np.int32)
    else:
        h = rng.standard_normal(n).astype(xdt_np)
    b.set_x(ph.NArray.from_numpy(h))
    return h


F32_PROGS = [lambda v, s: v + s, lambda v, s: v * s - v, lambda v, s: (v + 4) ** 2 * 0.5 - v, lambda v, s: v]


@pytest.mark.parametrize("mode,count", [(0, 400), (1, 400), (2, 400), (3, 40)])
def test_mixed_records_float32(ph, mode, count):
    rng = np.random.default_rng(201)
    recs = _records(rng, count, 2, 2)
    b = ph.GBatch(ph.DType.F32, ph.DType.F32, F32_PROGS, recs)
    x = _fill(ph, b, rng, np.float32)
    b.run(mode)
    _check(ph, b, recs, F32_PROGS, np.float32, x)


def test_modes_agree_and_graph_replays(ph):
    """Three graph replays with eager work in between, each equal to the
    direct launch; the unbatched form gives the same y, z and maxima."""
    rng = np.random.default_rng(203)
    recs = _records(rng, 300, 2, 2, max_side=60)
    b = ph.GBatch(ph.DType.F32, ph.DType.F32, F32_PROGS, recs)
    _fill(ph, b, rng, np.float32)
    b.run(1)
    y1, z1, m1, s1 = b.y.to_numpy(), b.z.to_numpy(), np.array(b.maxes()), np.array(b.sums())
    for rep in range(3):
        junk = ph.NArray.from_numpy(rng.standard_normal((64, 64)))
        (junk * 2.0).materialize()
        b.run(0)
        assert np.array_equal(b.y.to_numpy(), y1) and np.array_equal(b.z.to_numpy(), z1), rep
        assert np.array_equal(np.array(b.maxes()), m1, equal_nan=True), rep
        assert np.array_equal(np.array(b.sums()), s1, equal_nan=True), rep
    b.run(2)
    assert np.array_equal(b.y.to_numpy(), y1) and np.array_equal(b.z.to_numpy(), z1)
    assert np.array_equal(np.array(b.maxes()), m1, equal_nan=True)
    assert np.array_equal(np.array(b.sums()), s1, equal_nan=True)


def test_max_follows_enumerable(ph):
    """The `>` loop: a leading NaN is the result, a later NaN is skipped,
    the first of equal zeros gives the sign, -inf counts; arrays split over
    several work items (chunked) combine the same way."""
    cases = [
        [np.nan, 1.0, 2.0], [1.0, np.nan, 3.0, np.nan], [-0.0, 0.0, -1.0], [0.0, -0.0], [-np.inf, -np.inf],
        [-1.0, -0.0, 0.0, -0.0], [np.nan], [5.0, 5.0, -0.0], [-2.0] * 5000 + [0.0] + [-0.0] * 3 + [-3.0] * 30000,
        [-1.0] * 40000 + [-0.0, 0.0], [np.nan] + [1.0] * 50000, [1.0] * 50000 + [np.nan, 2.0],
    ]
    recs = [{"shape": [len(c)], "red": 0, "ops": ["max", "sum"]} for c in cases]
    b = ph.GBatch(ph.DType.F64, ph.DType.F64, [lambda v, s: v], recs)
    h = np.zeros(b.x.size)
    for k, c in enumerate(cases):
        h[b.offsets_xy[k]:b.offsets_xy[k] + len(c)] = c
    b.set_x(ph.NArray.from_numpy(h))
    for mode in (0, 2, 3):
        b.run(mode)
        mx = np.array(b.maxes())
        for k, c in enumerate(cases):
            em = O.max_seq(np.array(c))
            assert (np.isnan(em) and np.isnan(mx[k])) or np.float64(em).view(np.uint64) == mx[k].view(np.uint64), \
                (mode, k, em, mx[k])


def test_int32_and_float64_batches(ph):
    """Int32 x with wrapping and checked programs (y Int32), Float64 x with
    a comparison (y Bool) and reversed regions."""
    rng = np.random.default_rng(207)
    iprogs = [lambda v, s: v.wadd(s), lambda v, s: v // 3 + s, lambda v, s: v * 0.5 + s]
    recs = _records(rng, 200, 2, 1)
    b = ph.GBatch(ph.DType.I32, ph.DType.I32, iprogs, recs)
    x = _fill(ph, b, rng, np.int32)
    b.run(0)
    _check(ph, b, recs, iprogs, np.int32, x)
    fprogs = [lambda v, s: v > s, lambda v, s: v * v]
    recs = _records(rng, 200, 1, 1)
    b = ph.GBatch(ph.DType.F64, ph.DType.Bool, fprogs, recs)
    x = _fill(ph, b, rng, np.float64)
    b.run(1)
    _check(ph, b, recs, fprogs, np.float64, x)


def test_program_errors_raise(ph):
    """Checked Int32 arithmetic that overflows raises OverflowError, a
    division by a zero scalar DivisionByZeroError, in every mode."""
    recs = [{"shape": [100], "map": 0, "scalar": 2.0}, {"shape": [3, 4], "map": 0, "scalar": 1.0}]
    b = ph.GBatch(ph.DType.I32, ph.DType.I32, [lambda v, s: v * s], recs)
    h = np.zeros(b.x.size, np.int32)
    h[b.offsets_xy[0] + 57] = 2 ** 30
    b.set_x(ph.NArray.from_numpy(h))
    for mode in (0, 1, 2, 3):
        with pytest.raises(ph.OverflowError):
            b.run(mode)
    recs = [{"shape": [10], "map": 0, "scalar": 0.0}]
    b = ph.GBatch(ph.DType.I32, ph.DType.I32, [lambda v, s: v // s], recs)
    b.set_x(ph.NArray.from_numpy(np.arange(b.x.size, dtype=np.int32)))
    for mode in (1, 2, 3):
        with pytest.raises(ph.DivisionByZeroError):
            b.run(mode)


def test_creation_errors(ph):
    with pytest.raises(ph.IndexError):
        ph.GBatch(ph.DType.F32, ph.DType.F32, [], [{"shape": [5], "region": ["2..7"]}])
    with pytest.raises(ph.ArgumentError):
        ph.GBatch(ph.DType.F32, ph.DType.F32, [lambda v, s: v], [{"shape": [5], "map": 3}])
    with pytest.raises(ph.ArgumentError):  # a map result that is not y's dtype
        ph.GBatch(ph.DType.F32, ph.DType.F32, [lambda v, s: v * 2.0], [{"shape": [5], "map": 0}])
    with pytest.raises(ph.DimensionError):
        ph.GBatch(ph.DType.F32, ph.DType.F32, [], [{"shape": [2, 2, 2, 2]}])


def test_c5_workload_as_general_batch(ph):
    """The 20000-array C5 workload (y = x + s, x.sum, every axis by step 2)
    as a general batch: y and z equal the fixed batch's bit for bit, sums
    within the FP64 bound of the exact sums (checked on a sample)."""
    from phase.workloads import batched_shapes
    shapes, scalars, _ = batched_shapes(20000, 4)
    fixed = ph.Batch(shapes, scalars.tolist())
    fixed.fill_uniform(4)
    fixed.run(0)
    recs = [{"shape": list(s), "map": 0, "red": 1, "ops": "sum", "region": ["..2.."] * len(s),
             "scalar": float(scalars[k])} for k, s in enumerate(shapes)]
    b = ph.GBatch(ph.DType.F32, ph.DType.F32, [lambda v, s: v + s, lambda v, s: v], recs)
    xh = fixed.x.to_numpy()
    xg = np.zeros(b.x.size, np.float32)  # the same arrays in the general batch's layout
    for k, s in enumerate(shapes):
        n = int(np.prod(s))
        xg[b.offsets_xy[k]:b.offsets_xy[k] + n] = xh[fixed.offsets_x[k]:fixed.offsets_x[k] + n]
    b.set_x(ph.NArray.from_numpy(xg))
    b.tune()
    b.run(0)
    yf, zf = fixed.y.to_numpy(), fixed.z.to_numpy()
    yg, zg = b.y.to_numpy(), b.z.to_numpy()
    sums = b.sums()
    for k in range(0, len(shapes), 97):
        n = int(np.prod(shapes[k]))
        of, og = fixed.offsets_x[k], b.offsets_xy[k]
        assert np.array_equal(yg[og:og + n].view(np.uint32), yf[of:of + n].view(np.uint32)), k
        nz = int(np.prod([(d + 1) // 2 for d in shapes[k]]))
        assert np.array_equal(zg[b.offsets_z[k]:b.offsets_z[k] + nz].view(np.uint32),
                              zf[fixed.offsets_z[k]:fixed.offsets_z[k] + nz].view(np.uint32)), k
        xk = xh[of:of + n].astype(np.float64)
        ex = O.sum_exact(xk)
        assert abs(sums[k] - ex) <= n * 2.0 ** -52 * float(np.abs(xk).sum()), k


def test_int32_reductions_exact_and_checked(ph):
    """Int32 sums are exact and checked like Crystal's `+` (a running total
    outside Int32 raises OverflowError, even when the total ends inside it,
    across work items too); Int32 maxima exact."""
    rng = np.random.default_rng(211)
    progs = [lambda v, s: v, lambda v, s: v.wmul(s) - v]
    shapes = [[int(rng.integers(1, 60))] * int(rng.integers(1, 4)) for _ in range(150)] + [[70000], [3, 40000]]
    recs = [{"shape": sh, "red": k % 2, "ops": ["sum", "max"], "scalar": float(rng.integers(-3, 4))}
            for k, sh in enumerate(shapes)]
    b = ph.GBatch(ph.DType.I32, ph.DType.I32, progs, recs)
    h = rng.integers(-20000, 20000, b.x.size).astype(np.int32)
    b.set_x(ph.NArray.from_numpy(h))
    for mode in (0, 2):
        b.run(mode)
        sums, maxes = b.sums(), b.maxes()
        for k, rec in enumerate(recs):
            n = int(np.prod(rec["shape"]))
            xk = h[b.offsets_xy[k]:b.offsets_xy[k] + n].astype(np.int64)
            s = int(rec["scalar"])
            v = xk if rec["red"] == 0 else (xk * s).astype(np.int32).astype(np.int64) - xk
            assert sums[k] == float(v.sum()), (mode, k)
            assert maxes[k] == float(v.max()), (mode, k)
    big = 2 ** 31 - 1
    cases = [([big, -1, 1], False), ([big, 1, -5], True), ([2 ** 30, 2 ** 30, -2 ** 30, -2 ** 30], True),
             ([-2 ** 31, 0, -1], True), ([-2 ** 31, 5, -5], False), ([60000] * 40000, True),
             ([60000] * 35000 + [-60000] * 5000, False)]
    for vals, overflow in cases:
        recs = [{"shape": [len(vals)], "red": 0, "ops": "sum"}]
        b = ph.GBatch(ph.DType.I32, ph.DType.I32, [lambda v, s: v], recs)
        h = np.zeros(b.x.size, np.int32)
        h[:len(vals)] = vals
        b.set_x(ph.NArray.from_numpy(h))
        for mode in (0, 3):
            if overflow:
                with pytest.raises(ph.OverflowError):
                    b.run(mode)
            else:
                b.run(mode)
                assert b.sums()[0] == float(sum(vals)), (vals[:3], mode)
