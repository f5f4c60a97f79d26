"""Synthetic workload inputs shared by the benchmark and the tests.

Heat-equation initial field: 16 Gaussian bumps on an n x n grid, centres
U[0.1, 0.9) * n per axis, widths sigma U[0.01, 0.05) * n, amplitudes
U[0.5, 2.0), drawn in that order from numpy's PCG64 seeded with `seed`;
evaluated in Float64.

Batched uneven arrays: `n` Float32 arrays, rank uniform in {1, 2, 3}, each
side log-uniform in [8, 512] (floor of exp(U[ln 8, ln 513))), from numpy's
PCG64 generator seeded with `seed`, and one Float32 scalar per array, U[-1, 1).

As specified, E[side^3] ~ 1.08e7 makes the rank-3 third of 20000 arrays alone
~7.2e10 elements (~288 GB of Float32), all of one GPU's HBM before outputs.
The workload is therefore capped: an array with more than `cap` elements
(2^18 by default, 1 MiB of Float32) is drawn again (rank and sides), so every
array stays small, which is what the batched path is for.
"""
import math

import numpy as np


def batched_shapes(n=20000, seed=4, cap=1 << 18):
    """Returns (shapes, scalars, redraws): list of tuples, float32 array, and
    the number of arrays that were drawn again because of the cap."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(8.0), math.log(513.0)
    shapes = []
    redraws = 0
    for _ in range(n):
        while True:
            rank = int(rng.integers(1, 4))
            sides = tuple(int(math.floor(math.exp(v))) for v in rng.uniform(lo, hi, rank))
            if math.prod(sides) <= cap:
                break
            redraws += 1
        shapes.append(sides)
    scalars = rng.uniform(-1.0, 1.0, n).astype(np.float32)
    return shapes, scalars, redraws


def batch_stats(shapes):
    """Element total, slice total and a log2 histogram of array sizes."""
    sizes = [math.prod(s) for s in shapes]
    zs = [math.prod((d + 1) // 2 for d in s) for s in shapes]
    hist = {}
    for m in sizes:
        b = max(0, (m - 1).bit_length()) if m else 0
        hist[b] = hist.get(b, 0) + 1
    return {
        "arrays": len(shapes),
        "elements": int(sum(sizes)),
        "slice_elements": int(sum(zs)),
        "max_elements": int(max(sizes)) if sizes else 0,
        "log2_histogram": {f"<=2^{b}": hist[b] for b in sorted(hist)},
    }


def heat_initial_field(n=8192, seed=3, bumps=16):
    """u0[i, j] = sum_k A_k exp(-((i - cy_k)^2 + (j - cx_k)^2) / (2 sigma_k^2)), Float64."""
    rng = np.random.default_rng(seed)
    cy, cx = rng.uniform(0.1, 0.9, bumps) * n, rng.uniform(0.1, 0.9, bumps) * n
    sig = rng.uniform(0.01, 0.05, bumps) * n
    amp = rng.uniform(0.5, 2.0, bumps)
    yy = np.arange(n, dtype=np.float64)[:, None]
    xx = np.arange(n, dtype=np.float64)[None, :]
    u0 = np.zeros((n, n))
    for k in range(bumps):
        u0 += amp[k] * np.exp(-(((yy - cy[k]) ** 2) + ((xx - cx[k]) ** 2)) / (2 * sig[k] ** 2))
    return u0
