"""Pins the oracle's batched-array loop and its Float32 sequential sum.

No reference test covers batches (ph-core has no batched API), so the loop is
pinned piece by piece against independent routines: numpy's Float32 add,
numpy basic slicing x[::2, ::2, ::2] (the step-2 region on every axis), and
numpy's sequential Float32 cumulative sum; the Float32 sum also against
hand-computed absorption cases."""
import numpy as np
import pytest

import oracle_api as O
from phase.workloads import batch_stats, batched_shapes


def test_sum_seq_f32_known_answers(oracle):
    # 1e8 + 1 rounds back to 1e8 in Float32 (spacing 8): the sequence loses the 1
    assert O.sum_seq_f32(np.array([1e8, 1.0, -1e8], np.float32)) == 0.0
    # 2^24 + 1 + 1 + ... stays 2^24 (each +1 is a tie rounding to even)
    assert O.sum_seq_f32(np.array([2.0 ** 24] + [1.0] * 10, np.float32)) == 2.0 ** 24
    # reordered, the small terms survive
    assert O.sum_seq_f32(np.array([1.0] * 10 + [2.0 ** 24], np.float32)) == 2.0 ** 24 + 10
    assert O.sum_seq_f32(np.zeros(0, np.float32)) == 0.0


def test_sum_seq_f32_is_sequential_float32():
    rng = np.random.default_rng(1)
    for n in (1, 7, 1000, 65537):
        x = rng.uniform(-1, 1, n).astype(np.float32)
        assert O.sum_seq_f32(x) == np.cumsum(x, dtype=np.float32)[-1]


def test_batch_loop_matches_numpy():
    shapes, scalars, _ = batched_shapes(60, seed=4, cap=1 << 14)
    shapes += [(1,), (9,), (1, 1, 1), (2, 3), (5, 1, 7), (3, 3, 3)]
    scalars = np.concatenate([scalars, np.float32([0.5, -1.0, 0.25, 3.0, -0.125, 1e-3])])
    xo, zo, nx, nz = O.batch_layout(shapes)
    x = np.random.default_rng(2).uniform(-1, 1, nx).astype(np.float32)
    y, z, sums = O.batch_loop(shapes, scalars, x)
    for k, s in enumerate(shapes):
        n = int(np.prod(s))
        xk = x[xo[k]:xo[k] + n]
        assert np.array_equal(y[xo[k]:xo[k] + n], xk + scalars[k])
        zk = xk.reshape(s)[tuple(slice(None, None, 2) for _ in s)].reshape(-1)
        assert np.array_equal(z[zo[k]:zo[k] + zk.size], zk)
        assert sums[k] == np.cumsum(xk, dtype=np.float32)[-1]


def test_batch_loop_empty_axis_raises():
    with pytest.raises(O.OracleError) as e:
        O.batch_loop([(3, 0)], np.float32([1.0]), np.zeros(4, np.float32))
    assert e.value.code == -1  # IndexError


def test_workload_generator():
    shapes, scalars, redraws = batched_shapes()
    st = batch_stats(shapes)
    assert st["arrays"] == 20000 and len(scalars) == 20000
    assert st["max_elements"] <= 1 << 18
    assert all(1 <= len(s) <= 3 and all(8 <= d <= 512 for d in s) for s in shapes)
    assert redraws > 0
    again, sc2, _ = batched_shapes()
    assert again == shapes and np.array_equal(sc2, scalars)
