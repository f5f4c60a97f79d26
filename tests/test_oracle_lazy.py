"""Pins the oracle's lazily evaluated sources (or_set_chunk_lazy,
or_set_mask_lazy): the reference reads a source or mask element only when
its loop reaches it (src/n_array.cr:484-492 pulls `src.each` one element at a
time; :523-536 calls `mask.get(coord)` and `value.get(coord)` per
coordinate), so a source that is a view of the array being written sees the
loop's earlier writes.

Pins:
- closed forms of aliased assignments, derived by hand from the loop order
  (a transposed self-assignment, a shift that propagates, a reversal, a mask
  computed from the transposed array); evaluating the source before the
  writes (a snapshot) fails every one of them;
- without aliasing, the lazy forms equal the already pinned eager ones
  (or_set_mask, or_view_scatter, or_eval)."""
import random

import numpy as np

import oracle_api as O
from crystal import parse_literal
from test_oracle_pinning import _np_select, _random_chain, _random_literal


def P(lits):
    return [parse_literal(x) for x in lits]


def _tri_closed_form(z0):
    """z[all] = z.permute(1, 0) in lexicographic order: the strict upper
    triangle takes the old lower one (not written yet when read), the lower
    triangle reads back what the upper one just received (its own old
    value), the diagonal reads itself."""
    return np.tril(z0) + np.tril(z0, -1).T


def test_transposed_self_assignment_closed_form():
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 16):
        z0 = rng.integers(-50, 50, size=(n, n)).astype(np.int32)
        z = z0.copy()
        m = np.ones((n, n), np.bool_)
        O.set_mask_lazy(z, "x1", "x0", [(z, [("permute", [1, 0])]), (m, [])])
        assert np.array_equal(z, _tri_closed_form(z0)), n
        if n > 1:
            assert not np.array_equal(z, z0.T)  # the snapshot answer differs
        z = z0.copy()
        O.set_chunk_lazy(z, [], O.region_cover([n, n]), "x0", [(z, [("permute", [1, 0])])], [n, n])
        assert np.array_equal(z, _tri_closed_form(z0)), n


def test_shift_propagates():
    """z[1..] = z[..-2]: each element reads the one just written, so the first
    value runs through the whole array; with `+ 1` on the way it counts up."""
    rng = np.random.default_rng(1)
    for n in (2, 3, 17):
        z0 = rng.integers(-100, 100, size=n).astype(np.int32)
        src = O.region(P(["..-2"]), [n])
        dst = O.region(P(["1.."]), [n])
        z = z0.copy()
        O.set_chunk_lazy(z, [], dst, "x0", [(z, [("region", src)])], [n - 1])
        assert np.array_equal(z, np.full(n, z0[0], np.int32))
        z = z0.copy()
        O.set_chunk_lazy(z, [], dst, "(add x0 1i)", [(z, [("region", src)])], [n - 1])
        assert np.array_equal(z, z0[0] + np.arange(n, dtype=np.int32))
        # the other direction reads ahead of the writes: a plain shift
        z = z0.copy()
        O.set_chunk_lazy(z, [], src, "x0", [(z, [("region", dst)])], [n - 1])
        assert np.array_equal(z, np.concatenate([z0[1:], z0[-1:]]))


def test_reversal_closed_form():
    """z[..] = z.reverse: the first half takes the old second half, which the
    second half then reads back, so element i ends as z0[max(i, n-1-i)]."""
    for n in (1, 2, 7, 10):
        z0 = (np.arange(n, dtype=np.float64) * 1.5 - 3.0)
        z = z0.copy()
        O.set_chunk_lazy(z, [], O.region_cover([n]), "x0", [(z, [("reverse", [n])])], [n])
        i = np.arange(n)
        assert np.array_equal(z, z0[np.maximum(i, n - 1 - i)]), n


def test_mask_from_transposed_array_closed_form():
    """z[z.permute(1, 0) > 0] = 7 with the mask read lazily: (i, j) above the
    diagonal tests the old z[j, i]; below it, z[j, i] was already decided, so
    it tests old z[i, j] > 0 or old z[j, i] > 0; the diagonal tests itself."""
    rng = np.random.default_rng(2)
    for n in (1, 3, 12):
        z0 = rng.integers(-3, 3, size=(n, n)).astype(np.int32)
        z = z0.copy()
        O.set_mask_lazy(z, "(gt x0 0i)", 7, [(z, [("permute", [1, 0])])])
        up = np.triu(np.ones((n, n), bool), 1)
        lo = up.T
        pos, post = z0 > 0, z0.T > 0
        m = np.where(up, post, np.where(lo, pos | post, pos))
        assert np.array_equal(z, np.where(m, 7, z0)), n
        if n == 12:
            assert not np.array_equal(z, np.where(post, 7, z0))  # the snapshot answer


def test_lazy_forms_without_aliasing_equal_the_eager_ones():
    rng = random.Random(5)
    nrng = np.random.default_rng(5)
    for _ in range(120):
        rank = rng.randint(1, 3)
        shape = [rng.randint(1, 6) for _ in range(rank)]
        # masks: or_set_mask
        z0 = nrng.integers(-9, 9, size=shape).astype(np.int32)
        m = nrng.random(shape) < 0.5
        src = nrng.integers(100, 200, size=shape).astype(np.int32)
        a = z0.copy()
        O.set_mask_lazy(a, "x0", "x1", [(m, []), (src, [])])
        assert np.array_equal(a, O.set_mask(z0.copy(), m, src))
        a = z0.copy()
        O.set_mask_lazy(a, "x0", 42, [(m, [])])
        assert np.array_equal(a, O.set_mask(z0.copy(), m, 42))
        # chunks through a random MutableView chain: or_view_scatter of the
        # evaluated source
        chain, vshape, _ = _random_chain(rng, shape)
        lits = [_random_literal(rng, b) for b in vshape[: rng.randint(0, len(vshape))]]
        try:
            r = O.region(P(lits), vshape)
        except O.OracleError:
            continue
        if r.size == 0:
            continue
        other = nrng.integers(-1000, 1000, size=r.shape).astype(np.int32)
        exp_src, _ = O.eval_expr("(mul (add x0 3i) 2i)", [other])
        exp = O.view_scatter(z0.copy(), chain, r, exp_src)
        got = z0.copy()
        O.set_chunk_lazy(got, chain, r, "(mul (add x0 3i) 2i)", [(other, [])], r.shape)
        assert np.array_equal(got, exp), (shape, chain, lits)


def test_lazy_set_chunk_through_views_equals_numpy_when_disjoint():
    """A source view of the same array whose elements the region never
    writes reads only old values: equal to numpy's copy semantics."""
    rng = np.random.default_rng(7)
    z0 = rng.integers(0, 1000, size=(6, 8)).astype(np.int32)
    z = z0.copy()
    dst = O.region(P(["0..2..", ".."]), [6, 8])
    src = O.region(P(["1..2..", "..-1.."]), [6, 8])
    O.set_chunk_lazy(z, [], dst, "x0", [(z, [("region", src)])], dst.shape)
    exp = z0.copy()
    exp[0::2, :] = z0[1::2, ::-1]
    assert np.array_equal(z, exp)
    assert _np_select(dst) is not None
