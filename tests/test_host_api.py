"""Host side, no device needed: the C-ABI library's exports, the C++ mirror
of ph-core's range / region / coordinate API (against the specs' known
answers and the oracle), element-expression typing (against the oracle's
evaluator), and lowering of element programs to gfx950 code objects
(generated and compiled, never launched)."""
import ctypes
import json
import os
import random
import re

import numpy as np
import pytest

import oracle_api as O
from crystal import parse_literal, trace

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_spec_vectors.json")))
HEADER = os.path.join(ROOT, "include", "phhip.h")


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|long|const\s+char\s*\*)\s*(ph_\w+)\s*\(", text, flags=re.M)))


# ---------------------------------------------------------------- the C-ABI
def test_cabi_exports_every_declared_symbol(host):
    names = declared_functions()
    assert len(names) >= 30 and "ph_map" in names and "ph_batch_run" in names
    lib = ctypes.CDLL(os.path.join(host.LIB_DIR, "libphhip.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_cabi_without_device(host):
    if host.device_available():
        pytest.skip("a device is visible: the no-device error path is not reachable")
    lib = ctypes.CDLL(os.path.join(host.LIB_DIR, "libphhip.so"))
    lib.ph_last_error.restype = ctypes.c_char_p
    p = ctypes.c_void_p()
    assert lib.ph_alloc(ctypes.byref(p), ctypes.c_size_t(64)) == -11  # PH_E_NOINIT
    assert lib.ph_init(0) != 0
    assert lib.ph_last_error()  # names the failure


# ------------------------------------------------------------- range syntax
LITERALS = ["5..5", "..5", "...5", "5..0", "5..-1..", "0..2...10", "(5..-3)..1", "5..(-3..1)", "..", "...",
            "5..", "..-1..", "..-1...2", "(..-4)..", "..(-4..)", "9.step(by: -1, to: 0)",
            "0.step(by: 2, to: 10, exclusive: true)", "1..3..-1", "..-2..", "-3...-1", "0..2..", -1, 7, 0]


def test_range_literal_parse_matches_independent_parser(host):
    for lit in LITERALS:
        p = parse_literal(lit)
        r = host.RangeLiteral.parse(lit if isinstance(lit, str) else str(lit))
        if p[0] == "index":
            assert r.is_index and r.index == p[1], lit
        else:
            assert not r.is_index, lit
            assert (r.first, r.last, r.step, r.exclusive) == tuple(p[1:]), lit


def test_range_table_bound10(host):
    g = GOLD["ranges_valid_bound10"]
    for lit, f, s, l in g["cases"]:
        assert host.RangeSyntax.canonicalize_range(lit, g["bound"])[:3] == (f, s, l), lit
    for lit in GOLD["out_of_bounds_bound10"]["cases"] + GOLD["step_conflict"]["cases"]:
        with pytest.raises(IndexError):
            host.RangeSyntax.canonicalize_range(lit, 10)
    for lit in GOLD["empty_bound10"]["cases"]:
        assert host.RangeSyntax.canonicalize_range(lit, 10)[3] == 0


def _random_literal(rng, bound):
    if rng.random() < 0.15:
        return rng.randint(-bound - 1, bound)
    f = rng.choice([None, rng.randint(-bound - 1, bound + 1)])
    l = rng.choice([None, rng.randint(-bound - 1, bound + 1)])
    s = rng.choice([None, None, rng.choice([-3, -2, -1, 1, 2, 3, 4])])
    dots = "..." if rng.random() < 0.4 else ".."
    fs = "" if f is None else str(f)
    ls = "" if l is None else str(l)
    if s is None:
        return f"{fs}{dots}{ls}"
    return f"{fs}..{s}{dots}{ls}" if rng.random() < 0.7 else f"{fs}..({s}{dots}{ls})"


def test_canonicalize_matches_oracle_random(host):
    rng = random.Random(17)
    agree = raised = 0
    for _ in range(4000):
        bound = rng.randint(0, 12)
        lit = _random_literal(rng, bound)
        try:
            want = O.canonicalize_range(parse_literal(lit), bound)
        except O.OracleError as e:
            with pytest.raises(IndexError if e.code == O.E_INDEX else Exception):
                host.RangeSyntax.canonicalize_range(lit if isinstance(lit, str) else lit, bound)
            raised += 1
            continue
        assert host.RangeSyntax.canonicalize_range(lit, bound) == want, (lit, bound)
        agree += 1
    assert agree > 1000 and raised > 200


# ------------------------------------------------------------- IndexRegion
def test_index_region_known_answers(host):
    g = GOLD["region_2d"]
    r = host.IndexRegion(g["literal"], g["bound"])
    assert list(r.shape) == g["shape"] and list(r.first) == g["first"]
    assert list(r.last) == g["last"] and list(r.step) == g["step"]
    with pytest.raises(Exception):
        host.IndexRegion(g["bad_literal"], g["bound"])
    ra = host.IndexRegion.absolute(g["absolute_literal"])
    assert list(ra.shape) == g["shape"] and list(ra.first) == g["first"] and list(ra.last) == g["last"]

    g = GOLD["includes"]
    r = host.IndexRegion.absolute(g["literal"])
    for c in g["outside"]:
        assert not r.includes(c), c
    for a in g["inside_a"]:
        for b in g["inside_b"]:
            assert r.includes([a, b])

    g = GOLD["fits_in"]
    r = host.IndexRegion.absolute(g["literal"])
    assert r.fits_in(g["fits"]) and not r.fits_in(g["not_fits"])
    c = host.IndexRegion.cover(g["cover"]).trim(g["trim_to"])
    assert c.fits_in(g["trim_to"]) and not c.fits_in(g["not_fits"])

    g = GOLD["translate"]
    r = host.IndexRegion.absolute(g["literal"])
    probe = list(r.local_to_absolute(g["probe"]))
    t = host.IndexRegion.absolute(g["literal"]).translate(g["offset"])
    assert list(t.shape) == list(r.shape)
    assert list(t.first) == [a + b for a, b in zip(r.first, g["offset"])]
    assert list(t.local_to_absolute(g["probe"])) == [a + b for a, b in zip(probe, g["offset"])]
    for off in g["raise_offsets"]:
        with pytest.raises(IndexError):
            host.IndexRegion.absolute(g["literal"]).translate(off)

    g = GOLD["reverse"]
    r = host.IndexRegion.absolute(g["literal"])
    rv = host.IndexRegion.absolute(g["literal"]).reverse()
    assert list(rv.first) == g["reversed_first"] == list(r.last)
    assert list(rv.step) == [-s for s in r.step]

    for lit, f, s, l in GOLD["fully_defined_bound10"]["cases"]:
        r = host.IndexRegion.absolute([lit])
        assert (r.first[0], r.step[0], r.last[0]) == (f, s, l), lit
    for lit in GOLD["implicit_bounds_absolute_raise"]["cases"]:
        with pytest.raises(Exception):
            host.IndexRegion.absolute([lit])
    for lit in GOLD["negative_absolute_raise_index_error"]["cases"]:
        with pytest.raises(IndexError):
            host.IndexRegion.absolute([lit])


def test_index_region_matches_oracle_random(host):
    rng = random.Random(23)
    n = 0
    for _ in range(1500):
        rank = rng.randint(1, 3)
        bound = [rng.randint(1, 9) for _ in range(rank)]
        lits = [_random_literal(rng, b) for b in bound]
        drop = rng.random() < 0.5
        try:
            o = O.region([parse_literal(x) for x in lits], bound, drop)
        except O.OracleError:
            with pytest.raises(Exception):
                host.IndexRegion(lits, bound, drop)
            continue
        h = host.IndexRegion(lits, bound, drop)
        assert list(h.first) == o.get("first") and list(h.last) == o.get("last"), (lits, bound)
        assert list(h.step) == o.get("step") and list(h.proper_shape) == o.get("proper_shape")
        assert list(h.shape) == o.shape, (lits, bound, drop)
        if h.size:
            local = [rng.randrange(d) for d in h.proper_shape]
            assert list(h.local_to_absolute_unsafe(local)) == O.local_to_absolute(o, local)
        n += 1
    assert n > 300


def test_coord_and_shape_util(host):
    g = GOLD["coord_util"]
    for i, size, ok in g["has_index"]:
        assert host.CoordUtil.has_index(i, size) == ok, (i, size)
    for i, size, want in g["canonicalize_index"]:
        assert host.CoordUtil.canonicalize_index(i, size) == want
    for i, size in g["canonicalize_index_raise"]:
        with pytest.raises(IndexError):
            host.CoordUtil.canonicalize_index(i, size)
    g = GOLD["shape_util"]
    for a, b in g["compatible"]:
        assert host.ShapeUtil.compatible_shapes(a, b) and host.ShapeUtil.compatible_shapes(b, a)
    for a, b in g["incompatible"]:
        assert not host.ShapeUtil.compatible_shapes(a, b)


# -------------------------------------------------------- element programs
BLOCKS = [
    ("i32", lambda x: x + 4), ("i32", lambda x: x * 2 + 1), ("i32", lambda x: x // 3), ("i32", lambda x: x % 7),
    ("i32", lambda x: x / 2), ("i32", lambda x: x * 0.5), ("i32", lambda x: x ** 2), ("i32", lambda x: x.wadd(7)),
    ("f32", lambda x: (x + 4) ** 2 * 0.5 - x), ("f32", lambda x: x * 2), ("f32", lambda x: x > 0.5),
    ("f32", lambda x: x / 3), ("f64", lambda x: x // 2), ("f64", lambda x: x % 1.5), ("f64", lambda x: -x),
    ("i32", lambda x: x.cmp(3)), ("i32", lambda x: x & 12), ("i32", lambda x: ~x),
]
DT = {"i32": ("I32", np.int32), "f32": ("F32", np.float32), "f64": ("F64", np.float64)}
NP_OF = {"I32": np.int32, "F32": np.float32, "F64": np.float64, "Bool": np.bool_}


def test_expression_types_match_oracle(host):
    """Result type of each block under Crystal's promotion rules: the host
    expression's dtype equals the dtype the oracle's evaluator returns."""
    for dt, fn in BLOCKS:
        name, npdt = DT[dt]
        e = fn(host.placeholder(getattr(host.DType, name)))
        res, _ = O.eval_expr(trace(fn), [np.arange(1, 5).astype(npdt)])
        assert NP_OF[str(e.dtype).split(".")[-1]] == res.dtype.type, (dt, trace(fn), e.dtype, res.dtype)


def test_element_programs_compile_for_gfx950(host):
    """The fused map and the fused reduction kernel of a few blocks are
    generated and compiled by the runtime compiler (no launch, no device)."""
    for dt, fn in BLOCKS[:4] + BLOCKS[8:10]:
        name, _ = DT[dt]
        d = getattr(host.DType, name)
        e = fn(host.placeholder(d))
        host.compile_check(e, [d], 64, 96, 0)
        if str(e.dtype).split(".")[-1] != "Bool":
            host.compile_check(e, [d], 64, 1024, 1)
    two = host.input(0, host.DType.F64) + host.input(1, host.DType.F64)
    host.compile_check(two * host.input(2, host.DType.F64), [host.DType.F64] * 3, 16, 16, 0)


def test_tiled_map_programs_compile_for_gfx950(host):
    """The map that stages a transposed operand through LDS is generated
    and compiled for a few blocks and element types, mixed Int32 and
    Float32 operands among them (no launch, no device); the element index is
    refused there."""
    for dt, fn in BLOCKS[:3] + BLOCKS[8:10]:
        d = getattr(host.DType, DT[dt][0])
        host.compile_check(fn(host.placeholder(d)), [d], 64, 96, 2)
    two = host.input(0, host.DType.BF16) * host.input(1, host.DType.F32)
    host.compile_check(two, [host.DType.BF16, host.DType.F32], 300, 257, 2)
    I, F = host.DType.I32, host.DType.F32
    host.compile_check(host.input(0, I).wadd(host.input(1, I)), [I, I], 200, 132, 2)
    mixed = host.input(0, F) * host.input(1, I) - host.input(2, F)
    host.compile_check(mixed, [F, I, F], 64, 128, 2)
    with pytest.raises(host.ArgumentError):
        host.compile_check(host.input(0, host.DType.I32) + host.element_index(), [host.DType.I32], 64, 64, 2)


def test_ordered_map_programs_compile_for_gfx950(host):
    """The kernels that run a map or a copy in the reference's element order
    when an input reads the output's own buffer (classify, Jacobi pass,
    write-out, sequential loop; for copies the chain links, pointer jumping
    and gather) are generated and compiled (no launch, no device)."""
    for dt, fn in BLOCKS[:3] + BLOCKS[8:10]:
        d = getattr(host.DType, DT[dt][0])
        e = fn(host.placeholder(d))
        if e.dtype != d:
            continue
        host.compile_check(e, [d], 48, 48, 3)
    m = host.input(1, host.DType.Bool)
    sel = host.select(m, host.input(0, host.DType.I32) * 2, host.input(2, host.DType.I32))
    host.compile_check(sel, [host.DType.I32, host.DType.Bool, host.DType.I32], 40, 40, 3)
    for d in (host.DType.I32, host.DType.F64, host.DType.BF16, host.DType.Bool):
        host.compile_check(host.placeholder(d), [d], 30, 70, 4)


def test_reduction_prefold_compiles_for_gfx950(host):
    """Rows of more than 64 column tiles (2 x 2^20) and columns of more than
    64 row tiles (2^17 x 64) get the ordered pre-fold kernel besides the
    tile and the finish, for each state type."""
    for d in (host.DType.F32, host.DType.F64, host.DType.I32, host.DType.BF16):
        e = host.placeholder(d) * 2
        host.compile_check(e, [d], 2, 1 << 20, 1)
        host.compile_check(e, [d], 1 << 17, 64, 1)


def test_indexed_map_programs_compile_for_gfx950(host):
    """The kernels of a map whose source or destination is a view chain
    longer than one descriptor (operands addressed through index arrays:
    bounds check, write-inverse table, direct pass, and the ordered passes
    reading the inverse table) are generated and compiled."""
    for dt, fn in BLOCKS[:3] + BLOCKS[8:10]:
        d = getattr(host.DType, DT[dt][0])
        e = fn(host.placeholder(d))
        if e.dtype != d:
            continue
        host.compile_check(e, [d], 24, 16, 6)
    m = host.input(1, host.DType.Bool)
    sel = host.select(m, host.input(0, host.DType.I32) * 2, host.input(2, host.DType.I32))
    host.compile_check(sel, [host.DType.I32, host.DType.Bool, host.DType.I32], 40, 40, 6)


def test_element_index_programs(host):
    """A block that reads the element's lexicographic index (map_with_index,
    n_array.cr:605-611) compiles into the map kernel in its row-tiled and its
    flat form, is Int32-typed, and is refused by the fused reduction, which
    has no element order to give it."""
    i = host.element_index()
    assert str(i.dtype).endswith("I32")
    x = host.input(0, host.DType.F64)
    e = x * 2 + i
    assert str(e.dtype).endswith("F64")
    host.compile_check(e, [host.DType.F64], 64, 96, 0)     # flat: rows of fewer than 256
    host.compile_check(e, [host.DType.F64], 8, 3000, 0)    # row-tiled
    host.compile_check((i // 7) % 3, [host.DType.F64], 4, 4, 0)  # no input read at all
    with pytest.raises(host.ArgumentError, match="element index outside a map"):
        host.compile_check(e, [host.DType.F64], 64, 1024, 1)


def test_float_text_layout(host):
    """Float element text for JSON/YAML: shortest round-trip digits in
    Crystal's layout (parity unpinned: no reference spec prints a float; the
    cases follow the layout rule: plain notation while the decimal point is
    within 15 digits left / 3 zeros right of the first digit)."""
    cases = [(1.0, "1.0"), (0.1, "0.1"), (-2.5, "-2.5"), (100.0, "100.0"), (123.456, "123.456"),
             (1e14, "100000000000000.0"), (1e15, "1.0e+15"), (1.5e20, "1.5e+20"), (0.0001, "0.0001"),
             (1e-5, "1.0e-5"), (-0.0, "-0.0"), (0.0, "0.0"), (float("inf"), "Infinity"),
             (float("-inf"), "-Infinity"), (123456789012345.0, "123456789012345.0"),
             (0.30000000000000004, "0.30000000000000004")]
    for v, want in cases:
        assert host.float_to_s(v) == want, (v, host.float_to_s(v))
    assert host.float_to_s(float("nan")) == "NaN"
    assert host.float_to_s(float(np.float32(0.1)), True) == "0.1"
    rng = np.random.default_rng(3)
    for v in rng.standard_normal(2000) * 10.0 ** rng.integers(-30, 30, 2000):
        assert float(host.float_to_s(float(v)).replace("e+", "e")) == v  # reads back exactly


def test_general_batch_programs_compile_for_gfx950(host):
    """The general batch's persistent kernel is generated and compiled for
    programs over Int32, Float32 and Float64 elements (map into Int32,
    Float32, Float64 and Bool results; float results also as reductions);
    other input layouts are refused."""
    D = host.DType
    for xd in (D.F32, D.F64, D.I32):
        v, s = host.input(0, xd), host.input(1, xd)
        progs = [v + s, v * 2.0 - s, v > s]
        if xd == D.I32:
            progs += [v.wadd(s), v // s]
        for p in progs:
            host.compile_check(p, [xd, xd], 1, 1, 5)
    with pytest.raises(host.ArgumentError):
        host.compile_check(host.input(0, D.F32), [D.F32], 1, 1, 5)
    with pytest.raises(host.ArgumentError):
        host.compile_check(host.input(0, D.BF16) + host.input(1, D.BF16), [D.BF16, D.BF16], 1, 1, 5)
