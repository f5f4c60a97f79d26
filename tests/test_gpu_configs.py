"""The measured configurations at their full sizes.

C3: ProcView x -> (x + 4)**2 * 0.5 - x over 16384 x 16384 N(0, 1) data
(seed 2, Float32 and BF16 storage), sum and max in full and along both axes
in one pass.  On the device each tile sums in FP64 and tiles are combined as
compensated pairs of floats; each sum is compared with the oracle's exactly
rounded sum of the oracle-evaluated block (FP64 values, per Crystal's
promotion of Float32 * Float64), with a stated bound of 1e-12 relative;
maxima are order-free and must match bit for bit.

C4: 8192 x 8192 heat field of 16 Gaussian bumps (seed 3): three steps (one
two-step pass and one single step) are bit-exact against the oracle at full
size, and the 2000-step FP32 run stays within 1e-5 relative L2 of the FP64
run, with the FP64 energy non-increasing."""
import numpy as np
import pytest

import oracle_api as O
from crystal import trace
from phase.workloads import heat_initial_field

pytestmark = pytest.mark.gpu

BLOCK = lambda v: (v + 4) ** 2 * 0.5 - v  # noqa: E731


@pytest.mark.parametrize("storage", ["F32", "BF16"])
def test_c3_full_size_vs_fp64_oracle(ph, storage):
    n = 16384
    D = getattr(ph.DType, storage)
    x = ph.NArray.random([n, n], D, seed=2, dist="normal", a=0.0, b=1.0)
    res = x.process(BLOCK).reduce2d_all()
    h = x.to_numpy()
    f, flags = O.eval_expr(trace(BLOCK), [h], bf16=(0,) if storage == "BF16" else ())
    assert flags == 0 and f.dtype == np.float64
    del h, x
    full = O.sum_exact(f)
    rows = O.slices_reduce(f, 0, 2)  # exact, per row (slices(0))
    cols = O.slices_reduce(np.ascontiguousarray(f.T), 0, 2)  # exact, per column (slices(1))
    err = {
        "full_sum": abs(res["full_sum"] - full) / abs(full),
        "row_sum": float(np.max(np.abs(np.array(res["row_sum"]) - rows) / np.abs(rows))),
        "col_sum": float(np.max(np.abs(np.array(res["col_sum"]) - cols) / np.abs(cols))),
    }
    print(f"C3 {storage}: max relative error vs exact FP64 sums {err}")
    assert max(err.values()) <= 1e-12, err
    assert res["full_max"] == float(f.max())
    assert np.array_equal(np.array(res["row_max"]), f.max(axis=1))
    assert np.array_equal(np.array(res["col_max"]), f.max(axis=0))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_c4_full_size_steps_bit_exact(ph, dt):
    u0 = heat_initial_field(8192, seed=3).astype(dt)
    u = ph.NArray.from_numpy(u0)
    e = ph.heat2d(u, 0.2, 3, every=1)
    ref, eref = O.heat2d(u0, 0.2, 3, every=1)
    assert np.array_equal(u.to_numpy(), ref)
    assert np.allclose(e, eref, rtol=1e-13, atol=0)


def test_c4_full_run_fp32_vs_fp64(ph):
    u0 = heat_initial_field(8192, seed=3)
    u64 = ph.NArray.from_numpy(u0)
    u32 = ph.NArray.from_numpy(u0.astype(np.float32))
    e64 = np.array(ph.heat2d(u64, 0.2, 2000, every=100))
    e32 = np.array(ph.heat2d(u32, 0.2, 2000, every=100))
    a, b = u32.to_numpy().astype(np.float64), u64.to_numpy()
    rel = float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))
    print(f"C4: relative L2 of the FP32 run vs the FP64 run after 2000 steps: {rel:.3e}")
    assert rel <= 1e-5
    assert len(e64) == 21 and np.all(np.diff(e64) <= 1e-9 * e64[0])
    assert np.allclose(e32, e64, rtol=1e-5)
