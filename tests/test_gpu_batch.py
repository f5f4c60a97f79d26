"""Batched small arrays on the device (one persistent launch for the whole
batch) against the oracle's per-array host loop.

Bar: y = x + s and the step-2 slice copies are bit-exact (Float32 addition
is correctly rounded; a copy is a copy).  The per-array sums are FP32
compensated (Neumaier) sums; each is checked against the oracle's exact sum
with |gpu - exact| <= 2^-22 |exact| + n 2^-44 sum|x| (FP32 Neumaier bound:
one rounding of the result plus second-order terms), and all schedules must
give the same y and z."""
import numpy as np
import pytest

import oracle_api as O
from phase.workloads import batched_shapes

pytestmark = pytest.mark.gpu


def _host(b):
    return b.x.to_numpy(), b.y.to_numpy(), b.z.to_numpy(), np.array(b.sums())


def _check_sums(shapes, x, xoffs, sums):
    for k, s in enumerate(shapes):
        n = int(np.prod(s))
        xk = x[xoffs[k]:xoffs[k] + n].astype(np.float64)
        ex = O.sum_exact(xk)
        tol = 2.0 ** -22 * abs(ex) + n * 2.0 ** -44 * float(np.abs(xk).sum())
        assert abs(sums[k] - ex) <= tol, (k, s, sums[k], ex)


def _check_batch(b, shapes, scalars):
    x, y, z, sums = _host(b)
    xo = np.array(b.offsets_x)
    zo = np.array(b.offsets_z)
    ry, rz, _ = O.batch_loop(shapes, scalars, x, xo, zo)
    for k, s in enumerate(shapes):
        n = int(np.prod(s))
        a, e = xo[k], xo[k] + n
        assert np.array_equal(y[a:e].view(np.uint32), ry[a:e].view(np.uint32)), ("y", k, s)
    assert np.array_equal(z.view(np.uint32), rz.view(np.uint32))
    _check_sums(shapes, x, xo, sums)
    return y, z, sums


EDGE = [(1,), (7,), (64,), (65,), (1, 1, 1), (2, 2, 2), (8, 1, 9), (3, 5), (300, 301), (511, 3, 5),
        (17, 19, 23), (1, 40000), (40000, 1), (100000,), (9, 1), (1, 9)]


def _small_batch(n=300, seed=4):
    shapes, scalars, _ = batched_shapes(n, seed)
    shapes = shapes + EDGE
    rng = np.random.default_rng(11)
    scalars = np.concatenate([scalars, rng.uniform(-1, 1, len(EDGE)).astype(np.float32)])
    return shapes, scalars


def test_layout_matches_oracle(ph):
    shapes, scalars = _small_batch(50)
    b = ph.Batch(shapes, scalars.tolist())
    xo, zo, nx, nz = O.batch_layout(shapes)
    assert list(b.offsets_x) == xo.tolist() and list(b.offsets_z) == zo.tolist()
    assert b.x.size == nx and b.z.size == nz


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_batch_modes_vs_oracle(ph, mode):
    shapes, scalars = _small_batch()
    b = ph.Batch(shapes, scalars.tolist())
    b.fill_uniform(4)
    b.run(mode)
    _check_batch(b, shapes, scalars)


def test_batch_x_matches_oracle_random(ph, oracle):
    shapes, scalars = _small_batch(40)
    b = ph.Batch(shapes, scalars.tolist())
    b.fill_uniform(4)
    x = b.x.to_numpy()
    ref = oracle.fill_random((x.size,), 1, 4, 0, 0, -1.0, 1.0)
    assert np.array_equal(x.view(np.uint32), ref.view(np.uint32))


def test_batch_accessors(ph):
    shapes, scalars = _small_batch(20)
    b = ph.Batch(shapes, scalars.tolist())
    b.fill_uniform(7)
    b.run(0)
    x, y, z, _ = _host(b)
    for k in (0, 5, len(shapes) - 1):
        xk = b.x_at(k).to_numpy()
        assert xk.shape == tuple(shapes[k])
        assert np.array_equal(xk.reshape(-1), x[b.offsets_x[k]:b.offsets_x[k] + xk.size])
        assert np.array_equal(b.y_at(k).to_numpy(), (xk + np.float32(scalars[k])).astype(np.float32))
        sl = tuple(slice(None, None, 2) for _ in shapes[k])
        assert np.array_equal(b.z_at(k).to_numpy(), xk[sl])
    with pytest.raises(IndexError):
        b.x_at(len(shapes))


def test_graph_replays_track_inputs(ph):
    """Replays of the captured graph, with eager work and synchronisations
    between them, each match the oracle on the current contents of x."""
    shapes, scalars = _small_batch(120)
    b = ph.Batch(shapes, scalars.tolist())
    ref_sums = None
    for rep, seed in enumerate([4, 5, 4, 6]):
        b.fill_uniform(seed)
        junk = ph.NArray.random([1000, 1000], ph.DType.F32, 99) + 1.0  # eager work between replays
        _ = junk.sum()
        b.run(0)
        _, _, sums = _check_batch(b, shapes, scalars)
        if seed == 4:
            if ref_sums is None:
                ref_sums = sums
            else:
                assert np.array_equal(ref_sums, sums)  # deterministic


def test_schedules_agree(ph):
    """Every schedule (work-item size, grid, queue or static) gives the same y
    and z bit for bit and passes the sum bound; for a fixed work-item size
    the sums do not depend on grid or assignment."""
    shapes, scalars = _small_batch(200)
    b = ph.Batch(shapes, scalars.tolist())
    b.fill_uniform(4)
    base = None
    per_chunk = {}
    for chunk in (64, 2048, 32768):
        for grid in (1, 37, 1024):
            for dyn in (0, 1):
                b.schedule(chunk, grid, dyn)
                b.run(0)
                y, z, sums = _check_batch(b, shapes, scalars)
                if base is None:
                    base = (y, z)
                assert np.array_equal(base[0].view(np.uint32), y.view(np.uint32))
                assert np.array_equal(base[1].view(np.uint32), z.view(np.uint32))
                if chunk in per_chunk:
                    assert np.array_equal(per_chunk[chunk], sums)
                else:
                    per_chunk[chunk] = sums


def test_tune_then_run(ph):
    shapes, scalars = _small_batch(300)
    b = ph.Batch(shapes, scalars.tolist())
    b.fill_uniform(4)
    chunk, grid, dyn = b.tune()
    assert chunk >= 64 and grid >= 1 and dyn in (0, 1)
    b.fill_uniform(4)
    b.run(0)
    _check_batch(b, shapes, scalars)
    assert b.tune() == (chunk, grid, dyn)  # cached by size histogram


def test_batch_errors(ph):
    with pytest.raises(IndexError):
        ph.Batch([(3, 0, 2)], [1.0])
    with pytest.raises(ph.DimensionError):
        ph.Batch([(2, 2, 2, 2)], [1.0])
    with pytest.raises(ph.ArgumentError):
        ph.Batch([(2,)], [1.0, 2.0])
    b = ph.Batch([], [])
    b.run(0)
    b.run(2)
    with pytest.raises(ph.ArgumentError):
        b.schedule(8, 4, 1)


def test_full_workload_properties(ph):
    """The 20000-array workload: y - x == s exactly for every element, every
    slice equals x[::2, ::2, ::2] of its array, and sums meet the bound for
    every 97th array."""
    shapes, scalars, _ = batched_shapes()
    b = ph.Batch(shapes, scalars.tolist())
    b.fill_uniform(4)
    b.tune()
    b.run(0)
    x, y, z, sums = _host(b)
    xo, zo = np.array(b.offsets_x), np.array(b.offsets_z)
    for k, s in enumerate(shapes):
        n = int(np.prod(s))
        xk = x[xo[k]:xo[k] + n]
        assert np.array_equal(y[xo[k]:xo[k] + n], xk + scalars[k]), k
        sl = tuple(slice(None, None, 2) for _ in s)
        zk = xk.reshape(s)[sl]
        assert np.array_equal(z[zo[k]:zo[k] + zk.size], zk.reshape(-1)), k
    sample = list(range(0, len(shapes), 97))
    _check_sums([shapes[k] for k in sample], np.concatenate([x[xo[k]:xo[k] + int(np.prod(shapes[k]))] for k in sample]),
                np.cumsum([0] + [int(np.prod(shapes[k])) for k in sample])[:-1], sums[sample])
