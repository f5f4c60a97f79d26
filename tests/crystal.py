"""Test-side helpers written independently of the product:

* parse_literal: Crystal range-literal source text -> the tuple
  RangeSyntax.parse_range returns (src/range_syntax/range_syntax.cr:41-69),
  fed to the oracle.
* Tracer: records an element block (a Python lambda) as the oracle's prefix
  expression text, so one lambda drives both the device path (phase.Expr) and
  the oracle.
"""
import re

_STEP = re.compile(r"^\(?\s*(-?\d+)\s*\)?\.step\(by:\s*(-?\d+)\s*,\s*to:\s*(-?\d+)\s*(?:,\s*exclusive:\s*(true|false))?\s*\)$")


def _tokens(s):
    out = []
    i = 0
    while i < len(s):
        c = s[i]
        if c.isspace():
            i += 1
        elif s.startswith("...", i):
            out.append("...")
            i += 3
        elif s.startswith("..", i):
            out.append("..")
            i += 2
        elif c in "()":
            out.append(c)
            i += 1
        elif c in "+-" or c.isdigit():
            j = i + 1
            while j < len(s) and (s[j].isdigit() or s[j] == "_"):
                j += 1
            out.append(int(s[i:j].replace("_", "")))
            i = j
        else:
            raise ValueError("bad literal " + s)
    return out


def _parse(toks, pos):
    # operand
    def operand(p):
        if p < len(toks) and toks[p] == "(":
            node, p = _parse(toks, p + 1)
            assert toks[p] == ")"
            return node, p + 1
        if p < len(toks) and isinstance(toks[p], int):
            return ("int", toks[p]), p + 1
        return ("nil",), p

    left, pos = operand(pos)
    while pos < len(toks) and toks[pos] in ("..", "..."):
        excl = toks[pos] == "..."
        right, pos = operand(pos + 1)
        left = ("range", left, right, excl)
    return left, pos


def parse_literal(lit):
    """int -> ('index', i); str -> ('range', first, last, step, exclusive)."""
    if isinstance(lit, int):
        return ("index", lit)
    m = _STEP.match(lit.strip())
    if m:
        return ("range", int(m.group(1)), int(m.group(3)), int(m.group(2)), m.group(4) == "true")
    toks = _tokens(lit)
    node, pos = _parse(toks, 0)
    if pos != len(toks):
        raise ValueError("trailing tokens in " + lit)
    if node[0] == "int":
        return ("index", node[1])

    def val(n):
        return n[1] if n[0] == "int" else None

    _, b, e, excl = node
    if b[0] == "range":
        return ("range", val(b[1]), val(e), val(b[2]), excl)
    if e[0] == "range":
        return ("range", val(b), val(e[2]), val(e[1]), e[3])
    return ("range", val(b), val(e), None, excl)


class Tracer:
    """Builds oracle prefix text: (op a b)."""

    def __init__(self, text):
        self.t = text

    @staticmethod
    def lit(v):
        if isinstance(v, Tracer):
            return v.t
        if isinstance(v, bool):
            return "true" if v else "false"
        if isinstance(v, int):
            return f"{v}i"
        if isinstance(v, float):
            return f"{v!r}d"
        # phase.Scalar (explicit literal type)
        dt = str(v.dtype)
        val = v.value()
        if dt.endswith("F32"):
            return f"{float(val)!r}f"
        if dt.endswith("F64"):
            return f"{float(val)!r}d"
        if dt.endswith("Bool"):
            return "true" if val else "false"
        return f"{int(val)}i"

    def _b(op, rev=False):
        def f(self, other):
            a, b = self.t, Tracer.lit(other)
            if rev:
                a, b = b, a
            return Tracer(f"({op} {a} {b})")
        return f

    __add__ = _b("add")
    __radd__ = _b("add", True)
    __sub__ = _b("sub")
    __rsub__ = _b("sub", True)
    __mul__ = _b("mul")
    __rmul__ = _b("mul", True)
    __truediv__ = _b("div")
    __rtruediv__ = _b("div", True)
    __floordiv__ = _b("fdiv")
    __rfloordiv__ = _b("fdiv", True)
    __mod__ = _b("mod")
    __rmod__ = _b("mod", True)
    __pow__ = _b("pow")
    __rpow__ = _b("pow", True)
    __and__ = _b("and")
    __or__ = _b("or")
    __xor__ = _b("xor")
    __gt__ = _b("gt")
    __lt__ = _b("lt")
    __ge__ = _b("ge")
    __le__ = _b("le")
    __eq__ = _b("eq")
    __ne__ = _b("ne")
    wadd = _b("wadd")
    wsub = _b("wsub")
    wmul = _b("wmul")
    wpow = _b("wpow")
    cmp = _b("cmp")
    __hash__ = None

    def __neg__(self):
        return Tracer(f"(neg {self.t})")

    def __invert__(self):
        return Tracer(f"(bnot {self.t})")

    def abs(self):
        return Tracer(f"(abs {self.t})")


def trace(fn, n=1):
    xs = [Tracer(f"x{k}") for k in range(n)]
    return Tracer.lit(fn(*xs))
