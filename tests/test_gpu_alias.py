"""Sources and masks that read the array being written (src/n_array.cr:484-492
pulls `src.each` lazily; :523-536 evaluates `mask.get(coord)` and
`value.get(coord)` per coordinate), on the device against the oracle's lazily
evaluated restatement of those loops (or_set_chunk_lazy, or_set_mask_lazy)
and against closed forms.  Bit-exact: the element order decides which values
are read, and the device must read the same ones."""
import random
import time

import numpy as np
import pytest

import oracle_api as O
from crystal import parse_literal, trace
from test_oracle_pinning import _random_literal

pytestmark = pytest.mark.gpu


def P(lits):
    return [parse_literal(x) for x in lits]


def _i32(rng, shape, lo=-1000, hi=1000):
    return rng.integers(lo, hi, shape).astype(np.int32)


@pytest.mark.parametrize("n", [128, 200])
def test_transposed_self_assignment(ph, n):
    """z[all] = z.permute(1, 0) (mask form and chunk form): the upper
    triangle takes the old lower one, which the lower triangle reads back."""
    rng = np.random.default_rng(n)
    h = _i32(rng, (n, n))
    exp = np.tril(h) + np.tril(h, -1).T
    z = ph.NArray.from_numpy(h)
    z[ph.NArray.from_numpy(np.ones((n, n), bool))] = z.view().permute(1, 0)
    assert np.array_equal(z.to_numpy(), exp)
    z = ph.NArray.from_numpy(h)
    z[[]] = z.view().permute(1, 0)
    assert np.array_equal(z.to_numpy(), exp)


@pytest.mark.parametrize("kind", ["i32", "f64"])
def test_mask_with_transposed_procview_source(ph, kind):
    """z[mask] = z.view.permute(1, 0).process { ... } at 128 x 128 and a
    ragged 130 x 130."""
    rng = np.random.default_rng(3)
    fn = (lambda v: v * 3 - 7) if kind == "i32" else (lambda v: (v + 4) ** 2 * 0.5 - v)
    for n in (128, 130):
        h = _i32(rng, (n, n)) if kind == "i32" else rng.standard_normal((n, n))
        mh = rng.random((n, n)) < 0.6
        z = ph.NArray.from_numpy(h)
        z[ph.NArray.from_numpy(mh)] = z.view().permute(1, 0).process(fn)
        ref = h.copy()
        O.set_mask_lazy(ref, "x1", trace(fn), [(ref, [("permute", [1, 0])]), (mh, [])])
        assert np.array_equal(z.to_numpy(), ref), (kind, n)
        snap = h.copy()
        O.set_mask(snap, mh, O.eval_expr(trace(fn), [np.ascontiguousarray(h.T)])[0])
        assert not np.array_equal(ref, snap)  # the order is visible in this case


def test_mask_computed_from_transposed_array(ph):
    """z[z.view.permute(1, 0).process { |v| v > 0 }] = 7: the mask reads
    elements the loop has already decided."""
    rng = np.random.default_rng(5)
    for n in (128, 257):
        h = _i32(rng, (n, n), -3, 3)
        z = ph.NArray.from_numpy(h)
        z[z.view().permute(1, 0).process(lambda v: v > 0)] = 7
        ref = h.copy()
        O.set_mask_lazy(ref, "(gt x0 0i)", 7, [(ref, [("permute", [1, 0])])])
        assert np.array_equal(z.to_numpy(), ref), n
        up = np.triu(np.ones((n, n), bool), 1)
        m = np.where(up, h.T > 0, np.where(up.T, (h > 0) | (h.T > 0), h > 0))
        assert np.array_equal(ref, np.where(m, 7, h))


def test_chunk_from_procview_of_another_region(ph):
    """z[r] = z.view(r2).process { ... } for regions that overlap in every
    way: reading ahead of the writes, behind them, and both."""
    rng = np.random.default_rng(7)
    cases = [
        ((150, 170), ["1..", ".."], ["..-2", ".."]),         # reads the row just written
        ((150, 170), ["..-2", ".."], ["1..", ".."]),         # reads ahead: old values
        ((150, 170), ["0..2..", "..-1.."], ["1..2..", ".."]),  # disjoint rows
        ((64, 64), ["..", ".."], ["..-1..", ".."]),          # rows reversed
        ((9, 300), ["..", "5.."], ["..", "..-6"]),           # a shift along the row
    ]
    fn = lambda v: (v ^ 21) - 3  # noqa: E731  (bounded change per link: chains do not overflow)
    for shape, r, r2 in cases:
        h = _i32(rng, shape, -10 ** 6, 10 ** 6)
        z = ph.NArray.from_numpy(h)
        z[r] = z.view(*r2).process(fn)
        ref = h.copy()
        dst = O.region(P(r), list(shape))
        src = O.region(P(r2), list(shape))
        O.set_chunk_lazy(ref, [], dst, trace(fn), [(ref, [("region", src)])], src.shape)
        assert np.array_equal(z.to_numpy(), ref), (shape, r, r2)


def test_shifts_that_propagate(ph):
    """z[1..] = z[..-2]: each element reads the one just written.  As a copy
    (pointer jumping along the chain) the first value fills the array; with
    `+ 1` (a map: Jacobi passes, then the sequential loop for the deep
    chain) it counts up."""
    rng = np.random.default_rng(11)
    h = _i32(rng, (100000,))
    z = ph.NArray.from_numpy(h)
    z[["1.."]] = z.view("..-2")
    assert np.array_equal(z.to_numpy(), np.full(100000, h[0], np.int32))
    h2 = _i32(rng, (512, 256))
    z = ph.NArray.from_numpy(h2)
    z[["1..", ".."]] = z.view("..-2", "..")
    assert np.array_equal(z.to_numpy(), np.broadcast_to(h2[0], (512, 256)))
    h3 = _i32(rng, (3000,))
    z = ph.NArray.from_numpy(h3)
    z[["1.."]] = z.view("..-2").process(lambda v: v + 1)
    assert np.array_equal(z.to_numpy(), h3[0] + np.arange(3000, dtype=np.int32))
    # the other direction only reads ahead: a plain shift
    z = ph.NArray.from_numpy(h3)
    z[["..-2"]] = z.view("1..").process(lambda v: v + 1)
    assert np.array_equal(z.to_numpy(), np.concatenate([h3[1:] + 1, h3[-1:]]))


@pytest.mark.parametrize("dtype", [np.int32, np.float64, np.bool_])
def test_reversal_onto_itself(ph, dtype):
    """z[..] = z.reverse: element i ends as z0[max(i, n - 1 - i)]."""
    for n in (1, 2, 1001, 4096):
        h = (np.arange(n) % 7 - 3).astype(dtype)
        z = ph.NArray.from_numpy(h)
        z[[]] = z.view().reverse()
        i = np.arange(n)
        assert np.array_equal(z.to_numpy(), h[np.maximum(i, n - 1 - i)]), (dtype, n)


def test_self_copies_without_conflicts_run_in_parallel(ph):
    """z[0..2.., ..] = z[1..2.., ..] (disjoint rows) and z[r] = z.view(r)
    (the identity) on 8192^2 Int32: no element reads another's destination,
    so they run as ordinary parallel copies (the classifying pass finds no
    conflict) and finish in milliseconds."""
    rng = np.random.default_rng(13)
    h = _i32(rng, (8192, 8192))
    z = ph.NArray.from_numpy(h)
    ph.sync()
    t0 = time.perf_counter()
    z[["0..2..", ".."]] = z.view("1..2..", "..")
    ph.sync()
    t_disjoint = time.perf_counter() - t0
    exp = h.copy()
    exp[0::2] = h[1::2]
    assert np.array_equal(z.to_numpy(), exp)
    t0 = time.perf_counter()
    z[["1..3..-1", "..-2.."]] = z.view("1..3..-1", "..-2..")
    ph.sync()
    t_identity = time.perf_counter() - t0
    assert np.array_equal(z.to_numpy(), exp)
    print(f"disjoint rows {t_disjoint * 1e3:.2f} ms, identity {t_identity * 1e3:.2f} ms")
    assert t_disjoint < 1.0 and t_identity < 1.0


def _chain_view(rng, view, vshape, chain, allow_reshape):
    """A random Region/Permute/Reverse(/Reshape) chain applied to `view`,
    mirrored in the oracle's chain (newest first)."""
    for _ in range(rng.randint(1, 3)):
        ops = ["region", "permute", "reverse"] + (["reshape"] if allow_reshape else [])
        op = rng.choice(ops)
        if op == "region":
            lits = [_random_literal(rng, b) for b in vshape]
            try:
                r = O.region(P(lits), vshape)
            except O.OracleError:
                continue
            if r.size == 0:
                continue
            view = view.mutable_view(*lits) if hasattr(view, "mutable_view") else view.view(*lits)
            chain.insert(0, ("region", r))
            vshape = r.shape
        elif op == "permute":
            pat = list(range(len(vshape)))
            rng.shuffle(pat)
            view = view.permute(pat)
            chain.insert(0, ("permute", pat))
            vshape = [vshape[p] for p in pat]
        elif op == "reverse":
            view = view.reverse()
            chain.insert(0, ("reverse", list(vshape)))
        else:
            n = int(np.prod(vshape))
            divs = [d for d in range(1, n + 1) if n % d == 0]
            d = rng.choice(divs)
            ns = [d, n // d]
            view = view.reshape(ns)
            chain.insert(0, ("reshape", list(vshape), ns))
            vshape = ns
    return view, list(vshape)


def test_random_aliased_chains_against_the_oracle(ph):
    """A MutableView chain of z written from a view (or a ProcView) of z
    through another chain, shapes chosen so both sides have the same
    element count; copies and maps, with and without reshapes."""
    rng = random.Random(23)
    nrng = np.random.default_rng(23)
    done = 0
    for trial in range(200):
        if done >= 40:
            break
        rank = rng.randint(1, 3)
        shape = [rng.randint(2, 9) for _ in range(rank)]
        h = _i32(nrng, shape)
        z = ph.NArray.from_numpy(h)
        wchain, rchain = [], []
        wv, wshape = _chain_view(rng, z.mutable_view(), list(shape), wchain, allow_reshape=False)
        rv, rshape = _chain_view(rng, z.view(), list(shape), rchain, allow_reshape=True)
        if int(np.prod(wshape)) != int(np.prod(rshape)) or not isinstance(wv, ph.MutableView):
            continue
        if rshape != wshape:
            rv = rv.reshape(wshape)
            rchain.insert(0, ("reshape", list(rshape), list(wshape)))
        use_map = trial % 2 == 1
        src = rv.process(lambda v: v - 5) if use_map else rv
        wv[[]] = src
        ref = h.copy()
        O.set_chunk_lazy(ref, wchain, O.region_cover(wshape), "(sub x0 5i)" if use_map else "x0",
                         [(ref, rchain)], wshape)
        assert np.array_equal(z.to_numpy(), ref), (shape, wchain, rchain, use_map)
        done += 1
    assert done >= 25


def test_recurrence_down_the_rows_runs_per_column(ph):
    """z[1.., ..] = z[..-2, ..].process { |v| v + 1 }: each column is a chain
    of n - 1 elements, each reading the one just written above it, so
    z[i, j] = z0[0, j] + i.  The chains are deeper than the relaxation
    passes reach, and each runs on a thread of its own (one thread for
    the whole array took seconds at 4096^2)."""
    rng = np.random.default_rng(41)
    for n, m in [(2048, 2048), (4096, 4096)]:
        h = _i32(rng, (n, m))
        z = ph.NArray.from_numpy(h)
        ph.sync()
        t0 = time.perf_counter()
        z[["1..", ".."]] = z.view("..-2", "..").process(lambda v: v + 1)
        ph.sync()
        dt = time.perf_counter() - t0
        exp = h[0][None, :].astype(np.int64) + np.arange(n)[:, None]
        assert np.array_equal(z.to_numpy(), exp.astype(np.int32)), n
        print(f"row recurrence {n}x{m}: {dt * 1e3:.1f} ms")
        assert dt < 2.0
    # a block that is not affine, against the oracle's loop
    h = _i32(rng, (300, 257), -10 ** 6, 10 ** 6)
    fn = lambda v: (v ^ 21) - 3  # noqa: E731
    z = ph.NArray.from_numpy(h)
    z[["1..", ".."]] = z.view("..-2", "..").process(fn)
    ref = h.copy()
    dst = O.region(P(["1..", ".."]), [300, 257])
    src = O.region(P(["..-2", ".."]), [300, 257])
    O.set_chunk_lazy(ref, [], dst, trace(fn), [(ref, [("region", src)])], src.shape)
    assert np.array_equal(z.to_numpy(), ref)


def test_broadcast_recurrence_forms_a_tree(ph):
    """z[1.., ..] = z[..-2, 0..0].tile(1, m).process { ... }: every element
    of row i reads z[i - 1, 0], so that element has m dependents (a tree,
    not chains); row i ends up f^i(z0[0, 0]).  The elements read by others
    (column 0) form one chain, walked by one thread, and one pass then
    computes the rest from their parents.  A block of 3 rows, 20000 x 3
    and 3000 x 1024 (the loop in order took about a second)."""
    rng = np.random.default_rng(43)
    fn = lambda v: (v ^ 21) - 3  # noqa: E731
    for n, m in [(40, 33), (3, 5), (20000, 3), (3000, 1024)]:
        h = _i32(rng, (n, m))
        z = ph.NArray.from_numpy(h)
        ph.sync()
        t0 = time.perf_counter()
        z[["1..", ".."]] = z.view("..-2", "0..0").tile(1, m).process(fn)
        ph.sync()
        dt = time.perf_counter() - t0
        exp = h.copy()
        v = int(h[0, 0])
        for i in range(1, n):
            v = (v ^ 21) - 3
            exp[i, :] = v
        assert np.array_equal(z.to_numpy(), exp), (n, m)
        print(f"broadcast recurrence {n}x{m}: {dt * 1e3:.1f} ms")
    assert dt < 0.5


def test_tree_overflow_and_float_values(ph):
    """Trees run a pass per level: an Int32 `* 2` down the rows raises
    OverflowError once row 31 leaves Int32 (the reference's checked `*`),
    and stays exact below that; a Float64 tree is bit for bit the loop's."""
    n, m = 40, 33
    h = np.zeros((n, m), np.int32)
    h[0, 0] = 1
    z = ph.NArray.from_numpy(h)
    with pytest.raises(OverflowError):
        z[["1..", ".."]] = z.view("..-2", "0..0").tile(1, m).process(lambda v: v * 2)
    z = ph.NArray.from_numpy(h[:31].copy())
    z[["1..", ".."]] = z.view("..-2", "0..0").tile(1, m).process(lambda v: v * 2)
    exp = h[:31].copy()
    exp[1:, :] = (2 ** np.arange(1, 31))[:, None]
    assert np.array_equal(z.to_numpy(), exp)
    rng = np.random.default_rng(45)
    n, m = 200, 64
    hf = rng.standard_normal((n, m))
    zf = ph.NArray.from_numpy(hf)
    zf[["1..", ".."]] = zf.view("..-2", "0..0").tile(1, m).process(lambda v: v * 1.5 - 0.25)
    ref = hf.copy()
    v = float(hf[0, 0])
    for i in range(1, n):
        v = v * 1.5 - 0.25
        ref[i, :] = v
    got = zf.to_numpy()
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_tree_along_the_first_row(ph):
    """z[.., 1..] = z[0..0, ..-2].tile(n, 1).process { ... }: element
    (i, j) reads z[0, j - 1], which element (0, j - 1) wrote, so the
    elements of row 0 form a chain and every column hangs off it (a tree
    with a level per column), against the loop itself."""
    rng = np.random.default_rng(44)
    n, m = 100, 300
    h = _i32(rng, (n, m))
    fn = lambda v: (v * 3 + 7) % 1000  # noqa: E731
    z = ph.NArray.from_numpy(h)
    z[["..", "1.."]] = z.view("0..0", "..-2").tile(n, 1).process(fn)
    ref = [[int(v) for v in row] for row in h]
    for i in range(n):
        for j in range(m - 1):
            ref[i][j + 1] = (ref[0][j] * 3 + 7) % 1000
    assert np.array_equal(z.to_numpy(), np.array(ref, np.int32))
